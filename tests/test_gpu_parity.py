"""HIP tick vs the oracle on identical inputs: bit-exact on every byte of the segment and
synapse records, on spike counts and on the delay ring (integer, index and floating-point
state alike - the device reproduces the reference's arithmetic operation by operation)."""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _gpu_run(net, ticks, stim_fn):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net) == 0
    last = None
    for t in range(ticks):
        st = stim_fn(t)
        if st is not last:
            sim.upload_stimuli(st)
            last = st
        sim.run_tick(DT)
    out = sim.bion_data()
    res = (out.segments.copy(), out.synapses.copy(), sim.spikes(len(net.segments)),
           sim.delay_slots(len(net.synapses)), sim.stats())
    sim.close()
    return res


def _compare(net, ticks, oracle_stim_fn, gpu_stim_fn):
    o = H.OracleSim(net.copy())
    for t in range(ticks):
        o.step(DT, oracle_stim_fn(t))
    gseg, gsyn, gsp, gdl, stats = _gpu_run(net.copy(), ticks, gpu_stim_fn)
    print(f"tiles={stats.num_tiles} unsupported={stats.unsupported_modifiers}/{stats.unsupported_synapses} "
          f"hazard={stats.order_hazard_synapses}", flush=True)
    assert H.first_mismatch(o.net.segments, gseg) is None, H.first_mismatch(o.net.segments, gseg)
    assert H.first_mismatch(o.net.synapses, gsyn) is None, H.first_mismatch(o.net.synapses, gsyn)
    np.testing.assert_array_equal(o.spikes(), gsp)
    assert np.array_equal(o.delay_slots(), gdl)
    return o


@pytest.mark.parametrize("ticks", [1, 40, 400])
def test_config1_bitexact(ticks):
    net = bion.config1()
    clamp = H.soma_clamp(net, 200.0)
    ost = lambda t: H.Stimuli(current=clamp)
    gst_obj = backend.make_stimuli(current=clamp)
    o = _compare(net, ticks, ost, lambda t: gst_obj)
    if ticks >= 400:
        assert o.spikes().sum() > 0


def test_config2_small_bitexact():
    net = bion.config2(n_neurons=24, syn_per_neuron=100.0)
    pulses = bion.poisson_soma_pulses(net, 200.0, 300, DT, 300.0, 5)
    clamp = H.soma_clamp(net, 150.0)
    cur = [np.concatenate([clamp, p]) for p in pulses]
    gst = [backend.make_stimuli(current=c) for c in cur]
    o = _compare(net, 300, lambda t: H.Stimuli(current=cur[t]), lambda t: gst[t])
    assert o.spikes().sum() > 0
