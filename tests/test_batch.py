"""Batches, captured tick sequences and launch schedules: none of them may change a result.

- a batch of uneven networks (chain networks with their own parameter sets, a 12-channel tree
  network, passive cables) advanced together equals each network advanced alone, bit for bit,
  including per-network stimuli beyond the 256-entry upload limit and one-shot injections;
- ticks replayed from a captured hipGraph equal eager ticks, checked after every replay call
  with eager work and synchronisations in between;
- an autotuned or hand-set launch schedule gives the default schedule's results.
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _networks():
    nets = [bion.config4_network(k, n) for k, n in enumerate([6, 37, 13, 70])]
    nets.append(bion.config2(n_neurons=6, syn_per_neuron=40.0, seed=41))
    nets.append(bion.config3(n_cables=3, seg_per_cable=30, seed=42))
    nets.append(bion.config5_cell(44, n_seg=200))   # one 256-lane tile
    return nets


def _stimuli(k, net):
    cur = H.soma_clamp(net, 150.0 + 10.0 * k)
    ramp = np.zeros(1, bion.RAMP_DT)
    ramp["segment_id"], ramp["start_pA"], ramp["end_pA"], ramp["start_tick"], ramp["end_tick"] = 1, 0.0, 60.0, 3, 30
    ion = np.zeros(1, bion.ION_INJ_DT)
    ion["segment_id"], ion["ion_idx"], ion["amount_mM"] = 2, 2, 0.003
    return dict(current=cur, ramp=ramp, ion=ion if k % 2 == 0 else None)


def _alone(net, stim, chunks):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.upload_stimuli(backend.make_stimuli(**stim))   # the 256-per-kind upload path
    for c in chunks:
        for _ in range(c):
            sim.run_tick(DT)
    out = sim.bion_data()
    res = out.segments.copy(), out.synapses.copy(), sim.spikes(len(net.segments))
    sim.close()
    return res


def test_batch_equals_separate_networks():
    nets = _networks()
    chunks = [7, 16, 21]
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([n.copy() for n in nets]) == 0
    assert sim.network_count() == len(nets)
    for k, n in enumerate(nets):
        sim.set_stimuli(k, **_stimuli(k, n))
    for c in chunks:
        sim.run_ticks(DT, c)
    for k, n in enumerate(nets):
        seg, syn, spk = sim.batch_read(k)
        aseg, asyn, aspk = _alone(n, _stimuli(k, n), chunks)
        print(f"network {k}: {len(n.segments)} segments, spikes {int(spk.sum())}", flush=True)
        assert H.first_mismatch(aseg, seg) is None, (k, H.first_mismatch(aseg, seg))
        assert H.first_mismatch(asyn, syn) is None, (k, H.first_mismatch(asyn, syn))
        np.testing.assert_array_equal(aspk, spk)
    sim.close()


def test_more_stimuli_than_one_upload_holds():
    """600 current pulses on one network (the GpuStimulusData upload caps each kind at 256)."""
    net = bion.config4_network(9, 75)
    cur = H.soma_clamp(net, 180.0)
    extra = np.zeros(600 - len(cur), bion.CURRENT_STIM_DT)
    extra["segment_id"] = np.arange(len(extra)) % len(net.segments)
    extra["amplitude_pA"] = 5.0
    extra["duration_ms"] = 1e9
    cur = np.concatenate([cur, extra])
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.set_stimuli(0, current=cur)
    sim.run_ticks(DT, 30)
    out = sim.bion_data()
    o = H.OracleSim(net.copy())
    for _ in range(30):
        o.step(DT, H.Stimuli(current=cur))
    assert H.first_mismatch(o.net.segments, out.segments) is None, H.first_mismatch(o.net.segments, out.segments)
    assert H.first_mismatch(o.net.synapses, out.synapses) is None
    sim.close()


def test_graph_replays_equal_eager_ticks():
    net = bion.config2(n_neurons=8, syn_per_neuron=60.0, seed=43)
    clamp = H.soma_clamp(net, 170.0)
    eager = backend.Simulation(coherence=backend.COHERENCE_FAST)
    graph = backend.Simulation(coherence=backend.COHERENCE_FAST)
    for s in (eager, graph):
        assert s.load_network(net.copy()) == 0
        s.set_stimuli(0, current=clamp)
    eager.set_graphs(False)
    for c in [16, 16, 5, 32, 16]:
        for _ in range(c):
            eager.run_tick(DT)
        graph.run_ticks(DT, c)
        a, b = eager.bion_data(), graph.bion_data()
        assert eager.tick_count() == graph.tick_count()
        assert H.first_mismatch(a.segments, b.segments) is None, (c, H.first_mismatch(a.segments, b.segments))
        assert H.first_mismatch(a.synapses, b.synapses) is None
        np.testing.assert_array_equal(eager.spikes(len(net.segments)), graph.spikes(len(net.segments)))
    eager.close()
    graph.close()


def test_schedules_do_not_change_results():
    nets = _networks()[:3]
    sims = [backend.Simulation(coherence=backend.COHERENCE_FAST) for _ in range(3)]
    for s in sims:
        assert s.load_batch([n.copy() for n in nets]) == 0
        for k, n in enumerate(nets):
            s.set_stimuli(k, current=H.soma_clamp(n, 160.0))
    ran = sims[1].autotune(DT, 2)
    assert ran > 0
    sims[2].set_schedule(modifiers_block=64, diffusion_tiles=4, pre_grid=1024, plast_block=64, rrp_block=128,
                         step1_chunks=4)
    sims[0].run_ticks(DT, ran + 20)
    sims[1].run_ticks(DT, 20)
    sims[2].run_ticks(DT, ran + 20)
    print("autotuned schedule:", sims[1].schedule().as_dict(), flush=True)
    ref = [sims[0].batch_read(k) for k in range(3)]
    for s in sims[1:]:
        for k in range(3):
            seg, syn, spk = s.batch_read(k)
            assert H.first_mismatch(ref[k][0], seg) is None
            assert H.first_mismatch(ref[k][1], syn) is None
            np.testing.assert_array_equal(ref[k][2], spk)
    for s in sims:
        s.close()


def test_config4_batch_against_oracle():
    """A batch of the batched sweep's networks (chain neurons, each network its own seed and
    per-type g_max scaling, uneven sizes from 60 to 1400 segments) after autotuning and graph
    replay, each network compared with the oracle stepping it alone."""
    sizes = [9, 175, 48, 120, 30]
    nets = [bion.config4_network(k, n) for k, n in enumerate(sizes)]
    clamps = [bion.soma_current(n, 200.0) for n in nets]
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([n.copy() for n in nets]) == 0
    for k, c in enumerate(clamps):
        sim.set_stimuli(k, current=c)
    ran = sim.autotune(DT, 2)
    sim.run_ticks(DT, 300 - ran)
    for k, net in enumerate(nets):
        o = H.OracleSim(net.copy())
        for _ in range(300):
            o.step(DT, H.Stimuli(current=clamps[k]))
        seg, syn, spk = sim.batch_read(k)
        print(f"network {k}: {len(net.segments)} segments, spikes {int(spk.sum())}", flush=True)
        assert H.first_mismatch(o.net.segments, seg) is None, (k, H.first_mismatch(o.net.segments, seg))
        assert H.first_mismatch(o.net.synapses, syn) is None, (k, H.first_mismatch(o.net.synapses, syn))
        np.testing.assert_array_equal(o.spikes(), spk)
        assert spk.sum() > 0
    sim.close()
