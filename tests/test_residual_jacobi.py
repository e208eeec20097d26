"""Residual-terminated Jacobi on the GPU (an extension: each cell stops sweeping after the first
sweep whose largest update is <= tol, at most max sweeps) and the error of the reference's mixed
precision against the FP64 restatement.

- GPU residual mode equals the oracle's residual mode bit for bit, with one cell per tile
  (100-segment cables, 128-lane tiles) and several cells per tile (8-segment chains, 64-lane);
- after 500 ticks of 100-segment cables with linear +-20% ramps (tol 1e-6 mV, <= 200 sweeps),
  every field's relative error against FP64 is below its stated bound (printed per field).
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import fp64
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025
TOL, MAXS = 1e-6, 200
BOUND = {"V_m": 1e-4, "ATP": 1e-4, "temperature": 1e-6}   # intra ions 1e-6, extra 1e-9


@pytest.mark.parametrize("which", ["cables", "chains"])
def test_gpu_residual_mode_equals_oracle(which):
    net = (bion.config3(n_cables=6, seg_per_cable=100, seed=51) if which == "cables"
           else bion.config1(seed=52))
    stim = H.soma_clamp(net, 200.0)
    o = H.OracleSim(net.copy())
    fp64.lib().oracle_set_jacobi(o.ctx, TOL, MAXS)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.set_jacobi(TOL, MAXS)
    sim.set_stimuli(0, current=stim)
    for _ in range(150):
        o.step(DT, H.Stimuli(current=stim))
    sim.run_ticks(DT, 150)
    out = sim.bion_data()
    assert H.first_mismatch(o.net.segments, out.segments) is None, H.first_mismatch(o.net.segments, out.segments)
    assert H.first_mismatch(o.net.synapses, out.synapses) is None
    sim.close()


def test_mixed_precision_error_against_fp64():
    net = bion.config3(n_cables=8, seg_per_cable=100, seed=53)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.set_jacobi(TOL, MAXS)
    sim.run_ticks(DT, 500)
    seg = sim.bion_data().segments.copy()
    sim.close()
    ref = fp64.fp64_run(net, 500, DT, TOL, MAXS)
    err = fp64.field_errors(seg, ref)
    print("per-field max relative error after 500 ticks:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    for f, b in BOUND.items():
        assert err[f] < b, (f, err[f])
    assert max(err[f"intra[{k}]"] for k in range(7)) < 1e-6
    assert max(err[f"extra[{k}]"] for k in range(7)) < 1e-9
