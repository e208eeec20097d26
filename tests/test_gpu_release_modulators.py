"""Presynaptic autoreceptors (the reference's twelve built-in release modulators) and eCB
retrograde deposits on the device against the oracle, bit for bit on every byte of the segment
and synapse records (autoreceptor states included), spike counts and the delay ring:

- the modulated networks with both, and with each alone;
- immediate releases behind autoreceptors, some completed by the ordered pass (their vesicle
  recovery reads a presynaptic Ca that earlier synapses change);
- a batch of modulated networks, ticks replayed from captured hipGraphs;
- COMPAT coherence, with host edits of autoreceptor and eCB states between ticks."""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H
from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu
DT = 0.025


@pytest.mark.parametrize("seed,ar,ecb", [(1, True, True), (2, True, True), (3, True, False), (4, False, True)])
def test_modulated_network_bitexact(seed, ar, ecb):
    net = bion.modulated_network(seed, autoreceptors=ar, ecb=ecb)
    clamp = H.soma_clamp(net, 200.0)
    g = backend.make_stimuli(current=clamp)
    o = _compare(net, 1000, lambda t: H.Stimuli(current=clamp), lambda t: g)
    assert o.spikes().sum() > 0
    y = o.net.synapses
    if ar:
        ar0 = net.synapses["pre_autoreceptors"]["state"][:, :, 0]
        assert (y["pre_autoreceptors"]["state"][:, :, 0] != ar0).sum() > 100
    if ecb:
        assert (y["cleft_neurotransmitter"][:, 9] > 0).sum() > 10


def test_immediate_releases_behind_autoreceptors():
    net = bion.modulated_network(8, n_neurons=12)
    rng = np.random.default_rng(8)
    y = net.synapses
    onto_soma = rng.random(len(y)) < 0.5   # post chains on the somas, the presynaptic segments
    y["post_segment"][onto_soma] = rng.integers(0, 12, int(onto_soma.sum())) * 6
    now = rng.random(len(y)) < 0.4
    y["delay"][now] = 0.0
    clamp = H.soma_clamp(net, 220.0)
    g = backend.make_stimuli(current=clamp)
    o = _compare(net, 800, lambda t: H.Stimuli(current=clamp), lambda t: g)
    assert o.spikes().sum() > 0


def test_batch_of_modulated_networks():
    nets = [bion.modulated_network(s, n_neurons=n) for s, n in ((11, 10), (12, 16), (13, 7))]
    clamps = [H.soma_clamp(n, 190.0 + 10.0 * k) for k, n in enumerate(nets)]
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([n.copy() for n in nets]) == 0
    for k, c in enumerate(clamps):
        sim.set_stimuli(k, current=c)
    for c in (16, 100, 284):
        sim.run_ticks(DT, c)
    for k, net in enumerate(nets):
        o = H.OracleSim(net.copy())
        for _ in range(400):
            o.step(DT, H.Stimuli(current=clamps[k]))
        segs, syns, spk = sim.batch_read(k)
        assert H.first_mismatch(o.net.segments, segs) is None, (k, H.first_mismatch(o.net.segments, segs))
        assert H.first_mismatch(o.net.synapses, syns) is None, (k, H.first_mismatch(o.net.synapses, syns))
        np.testing.assert_array_equal(o.spikes(), spk)
    sim.close()


def test_compat_coherence_with_host_edits():
    net = bion.modulated_network(21, n_neurons=8)
    clamp = H.soma_clamp(net, 200.0)
    sim = backend.Simulation(coherence=backend.COHERENCE_COMPAT)
    assert sim.load_network(net.copy()) == 0
    sim.upload_stimuli(backend.make_stimuli(current=clamp))
    o = H.OracleSim(net.copy())
    has_ar = np.nonzero(net.synapses["pre_autoreceptor_count"] > 0)[0]
    ecb = np.nonzero(net.segments["modifiers"]["type"] == bion.ECB_SYNTHESIS)
    for t in range(300):
        if t in (50, 180):   # the same edits on both sides between ticks
            for recs in (sim.bion_data(), o.net):
                recs.synapses["pre_autoreceptors"]["state"][has_ar[:5], 0, 0] = 0.9
                recs.segments["modifiers"]["state"][ecb[0][:3], ecb[1][:3], 0] = 4.0
        sim.run_tick(DT)
        o.step(DT, H.Stimuli(current=clamp))
    out = sim.bion_data()
    assert H.first_mismatch(o.net.segments, out.segments) is None, H.first_mismatch(o.net.segments, out.segments)
    assert H.first_mismatch(o.net.synapses, out.synapses) is None, H.first_mismatch(o.net.synapses, out.synapses)
    sim.close()
