"""Transmitter uptake on the device against the oracle, bit for bit (the oracle equals the
compiled reference on these networks):

- glial capture: every chemical synapse of a network with GLAST segments hands part of its
  cleft glutamate, glycine and GABA to the glial segment nearest its post segment in depth
  (src/simulation_host.cpp:349-354, 466-490), merged into that segment's transmitters in
  synapse order, including glia that also receive synapses, depth ties and more than 256 glia;
- the six Na+-coupled transporters as step-1 modifiers (include/modifier_helpers.h:418-455);
- ionotropic receptors as segment modifiers, which read the segment's transmitters in step 1;
- per-network glia lists in a batch.
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _gpu(net, ticks, stim_fn):
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
    res = out.segments.copy(), out.synapses.copy(), sim.spikes(len(net.segments)), sim.stats()
    sim.close()
    return res


def _compare(net, ticks, ost, gst):
    o = H.OracleSim(net.copy())
    for t in range(ticks):
        o.step(DT, ost(t))
    seg, syn, spk, stats = _gpu(net.copy(), ticks, gst)
    print(f"unsupported={stats.unsupported_modifiers}/{stats.unsupported_synapses} "
          f"hazard={stats.order_hazard_synapses} spikes={int(spk.sum())}", flush=True)
    assert stats.unsupported_modifiers == 0 and stats.order_hazard_synapses == 0
    assert H.first_mismatch(o.net.segments, seg) is None, H.first_mismatch(o.net.segments, seg)
    assert H.first_mismatch(o.net.synapses, syn) is None, H.first_mismatch(o.net.synapses, syn)
    np.testing.assert_array_equal(o.spikes(), spk)
    return seg, syn, spk


@pytest.mark.parametrize("seed,n_glia,first", [(3, 5, False), (4, 6, True), (5, 270, False)])
def test_glial_capture_bitexact(seed, n_glia, first):
    net = bion.glia_network(seed, n_glia=n_glia, glia_first=first)
    cur = H.soma_clamp(net, 150.0)
    cur = cur[net.segments["modifiers"]["type"][cur["segment_id"], 0] != bion.GLAST]   # neurons only
    gst = backend.make_stimuli(current=cur)
    seg, syn, spk = _compare(net, 300, lambda t: H.Stimuli(current=cur), lambda t: gst)
    glia = np.nonzero(net.segments["modifiers"]["type"][:, 0] == bion.GLAST)[0]
    assert seg["neurotransmitter"][glia, 0].max() > 0.0 and spk.sum() > 0


def test_transporters_take_up_injected_transmitters():
    net = bion.glia_network(9, n_glia=4)
    net.synapses = net.synapses[:0]
    n = len(net.segments)
    lig = np.zeros(12, bion.LIG_INJ_DT)
    lig["segment_id"] = np.arange(12) * (n // 12)
    lig["segment_id"][-3:] = n - 3 + np.arange(3)
    lig["nt_idx"] = np.tile([0, 1, 7], 4)
    lig["amount"] = np.linspace(0.5, 30.0, 12)
    cur = H.soma_clamp(net, 150.0)
    g_on, g_off = backend.make_stimuli(current=cur, lig=lig), backend.make_stimuli(current=cur)
    seg, _, _ = _compare(net, 200, lambda t: H.Stimuli(current=cur, lig=lig if t % 40 == 2 else None),
                         lambda t: g_on if t % 40 == 2 else g_off)
    got = seg["neurotransmitter"][lig["segment_id"], lig["nt_idx"]]
    assert np.all(got > 0.0) and np.all(got < 5 * lig["amount"])


def test_ionotropic_receptors_as_segment_modifiers():
    net = bion.segment_receptor_network(13)
    cur = H.soma_clamp(net, 160.0)
    gst = backend.make_stimuli(current=cur)
    _compare(net, 300, lambda t: H.Stimuli(current=cur), lambda t: gst)


def test_glia_lists_are_per_network_in_a_batch():
    nets = [bion.glia_network(21, n_glia=3), bion.config1(seed=22), bion.glia_network(23, n_glia=4, glia_first=True)]
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([n.copy() for n in nets]) == 0
    clamps = [H.soma_clamp(n, 150.0) for n in nets]
    for k, c in enumerate(clamps):
        sim.set_stimuli(k, current=c)
    sim.run_ticks(DT, 120)
    for k, n in enumerate(nets):
        o = H.OracleSim(n.copy())
        for _ in range(120):
            o.step(DT, H.Stimuli(current=clamps[k]))
        seg, syn, spk = sim.batch_read(k)
        assert H.first_mismatch(o.net.segments, seg) is None, (k, H.first_mismatch(o.net.segments, seg))
        assert H.first_mismatch(o.net.synapses, syn) is None, (k, H.first_mismatch(o.net.synapses, syn))
        np.testing.assert_array_equal(o.spikes(), spk)
    sim.close()
