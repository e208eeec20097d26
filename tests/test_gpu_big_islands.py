"""Islands larger than one tile (more than 256 segments or 256 connections) on the GPU: the
one-workgroup-per-island electrodiffusion path against the oracle, byte for byte, for
- one ~300-segment island of ~1700 connections, glia coupled extracellularly to the neurons of
  three spinal segments as the reference's spinal network builder does
  (include/cpg_config.h:657-713), with synapses and, without, with transmitter diffusion;
- a 1000-compartment cell next to small neurons (big island and tiles in one launch sequence),
  in the fixed 50-sweep mode, the residual-terminated mode and cable mode with coupling scales.
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import fp64
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _run(net, ticks, stim_fn, setup=None, oracle_setup=None):
    o = H.OracleSim(net.copy())
    if oracle_setup:
        oracle_setup(o)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    assert sim.stats().big_islands >= 1
    if setup:
        setup(sim)
    last = None
    for t in range(ticks):
        kw, gst = stim_fn(t)
        o.step(DT, H.Stimuli(**kw))   # fresh each tick: the CPU tick consumes one-shot injections
        if gst is not last:
            sim.upload_stimuli(gst)
            last = gst
        sim.run_tick(DT)
    g = sim.bion_data()
    assert H.first_mismatch(o.net.segments, g.segments) is None, H.first_mismatch(o.net.segments, g.segments)
    assert H.first_mismatch(o.net.synapses, g.synapses) is None, H.first_mismatch(o.net.synapses, g.synapses)
    np.testing.assert_array_equal(o.spikes(), sim.spikes(len(net.segments)))
    sim.close()
    return o


@pytest.mark.parametrize("synapses", [True, False])
def test_glia_coupled_spinal_island(synapses):
    net = bion.spinal_glia_network(5, synapses=synapses)
    clamp = H.soma_clamp(net, 180.0)
    lig = np.zeros(2, bion.LIG_INJ_DT)
    lig["segment_id"], lig["nt_idx"], lig["amount"] = [5, 40], [0, 3], 0.5
    plain = (dict(current=clamp), backend.make_stimuli(current=clamp))
    inj = (dict(current=clamp, lig=lig), backend.make_stimuli(current=clamp, lig=lig))

    def stim(t):
        return inj if (not synapses and t % 10 == 2) else plain
    o = _run(net, 300, stim)
    assert o.spikes().sum() > 0


@pytest.mark.parametrize("mode", ["fixed", "residual", "cable"])
def test_thousand_compartment_cell(mode):
    net = bion.long_cell(6)
    clamp = H.soma_clamp(net, 150.0)
    pair = (dict(current=clamp), backend.make_stimuli(current=clamp))
    setup = oracle_setup = None
    if mode == "residual":
        setup = lambda s: s.set_jacobi(1e-6, 200)
        oracle_setup = lambda o: fp64.lib().oracle_set_jacobi(o.ctx, 1e-6, 200)
    if mode == "cable":
        scale = np.random.default_rng(2).uniform(0.5, 1.5, len(net.connections)).astype(np.float32)

        def setup(s):
            s.set_cable_mode(True)
            s.set_coupling_scale(scale)
        oracle_setup = lambda o: o.set_cable(True, scale)
    o = _run(net, 250, lambda t: pair, setup, oracle_setup)
    assert o.spikes().sum() > 0
