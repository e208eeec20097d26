"""The HIP tick against the committed golden fixtures over their whole length (written by
tests/golden/gen_golden.py from the reference's own compiled CPU library): every spike event
(tick, segment), every release event (tick, synapse), somatic V_m every 10 ticks, SHA-256 of all
segment and synapse records every 100 ticks and the final records - 4000 ticks of the 64-neuron
network and 1000 ticks of the 24-neuron 12-channel tree network. And the full-size 12-channel
network (100k neurons x 32 segments, 10M synapses) against the compiled reference."""
import hashlib
import os

import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DT = 0.025


def _gpu_replay(path, net, stims, ticks):
    g = np.load(path)
    assert str(g["net_hash"]) == hashlib.sha256(net.segments.tobytes() + net.synapses.tobytes()).hexdigest()
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    n = len(net.segments)
    somas = np.nonzero(net.segments["segment_type"] == 1)[0]
    prev = np.zeros(n, np.uint32)
    spikes, rels, vs, hs, hy = [], [], [], [], []
    last = None
    for t in range(ticks):
        st = stims(t)
        if st is not last:
            sim.upload_stimuli(st)
            last = st
        sim.run_tick(DT)
        sp = sim.spikes(n)
        spikes += [(t, int(i)) for i in np.nonzero(sp != prev)[0]]
        prev = sp
        b = sim.bion_data()
        rels += [(t, int(y)) for y in np.nonzero(b.synapses["refractory"] == 31)[0]]
        if t % 10 == 9:
            vs.append(b.segments["V_m"][somas].copy())
        if t % 100 == 99:
            hs.append(hashlib.sha256(b.segments.tobytes()).hexdigest())
            hy.append(hashlib.sha256(b.synapses.tobytes()).hexdigest())
    b = sim.bion_data()
    fs, fy = b.segments.copy(), b.synapses.copy()
    sim.close()
    np.testing.assert_array_equal(np.array(spikes, np.int64).reshape(-1, 2), g["spike_events"])
    np.testing.assert_array_equal(np.array(rels, np.int64).reshape(-1, 2), g["release_events"])
    assert np.array(vs).tobytes() == g["v_soma"].tobytes()
    assert list(hs) == list(g["seg_hash"]) and list(hy) == list(g["syn_hash"])
    assert H.first_mismatch(g["final_segments"], fs) is None
    assert H.first_mismatch(g["final_synapses"], fy) is None
    return g


def test_trace_config1_4000_ticks_on_gpu():
    net = bion.config1()
    st = backend.make_stimuli(current=H.soma_clamp(net, 200.0))
    g = _gpu_replay(os.path.join(GOLD, "trace_config1.npz"), net, lambda t: st, 4000)
    assert len(g["spike_events"]) > 1000 and len(g["release_events"]) > 100


def test_trace_config2_small_1000_ticks_on_gpu():
    net = bion.config2(n_neurons=24, syn_per_neuron=100.0)
    pulses = bion.poisson_soma_pulses(net, 200.0, 1000, DT, 300.0, 5)
    clamp = H.soma_clamp(net, 150.0)
    sts = [backend.make_stimuli(current=np.concatenate([clamp, p])) for p in pulses]
    g = _gpu_replay(os.path.join(GOLD, "trace_config2s.npz"), net, lambda t: sts[t], 1000)
    assert len(g["release_events"]) > 100


@pytest.mark.timeout(900)
def test_full_size_network():
    """100k neurons x 32 segments = 3.2M segments, 10M synapses (the benchmark network) with
    Poisson pulses, against the compiled reference (or, where it was not built, the oracle) over 1
    tick (BC_FULL_SIZE_TICKS ticks): V_m, every ion, ATP, I_buffer, spike counts, every synapse's vesicle pool,
    bouton Ca, refractory counter, weight, spine Ca and clefts in full; whole records byte for
    byte on a seeded sample of 200k segments and 500k synapses (32-bit slot offsets, tile packing
    and post-pass windows at full population)."""
    import time
    t0 = time.time()
    net = bion.config2(n_neurons=100_000)
    # a CPU tick takes about a minute at this size: 1 tick by default (BC_FULL_SIZE_TICKS=5 for
    # a longer check)
    ticks = int(os.environ.get("BC_FULL_SIZE_TICKS", "1"))
    pulses = bion.poisson_soma_pulses(net, 5.0, ticks, DT, 1000.0, 0xB10C0003)
    print(f"built {time.time() - t0:.1f}s", flush=True)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    for t in range(ticks):
        sim.upload_stimuli(backend.make_stimuli(current=pulses[t][:256]))
        sim.run_tick(DT)
    b = sim.bion_data()
    n, ns = len(net.segments), len(net.synapses)
    gs, gy, gsp = b.segments, b.synapses, sim.spikes(n)
    print(f"gpu done {time.time() - t0:.1f}s", flush=True)
    stim = lambda t: H.Stimuli(current=pulses[t][:256])
    v0 = net.segments["V_m"].copy()
    rsp = H.run_reference_inproc(net, ticks, DT, stim)
    if rsp is None:   # the oracle (about a minute per tick at this size)
        o = H.OracleSim(net)
        for t in range(ticks):
            o.step(DT, stim(t))
        rsp = o.spikes()
        assert np.array_equal(sim.delay_slots(ns), o.delay_slots())
    print(f"cpu done {time.time() - t0:.1f}s", flush=True)
    rs, ry = net.segments, net.synapses
    for f in ("V_m", "intra_ions", "extra_ions", "ATP_level", "I_buffer"):
        assert gs[f].tobytes() == rs[f].tobytes(), f
    for f in ("RRP", "bouton_Ca", "refractory", "weight", "spine_Ca", "cleft_neurotransmitter"):
        assert gy[f].tobytes() == ry[f].tobytes(), f
    np.testing.assert_array_equal(gsp, rsp)
    rng = np.random.default_rng(11)
    si = np.sort(rng.choice(n, 200_000, replace=False))
    yi = np.sort(rng.choice(ns, 500_000, replace=False))
    assert H.first_mismatch(rs[si], gs[si]) is None
    assert H.first_mismatch(ry[yi], gy[yi]) is None
    assert (rs["V_m"] != v0).sum() > n // 2   # the tick moved the state
    sim.close()
