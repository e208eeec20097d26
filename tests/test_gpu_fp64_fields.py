"""Each field's error of the HIP tick (the reference's mixed precision, reproduced bit for bit)
against the FP64 build of the oracle after 1000 ticks, on the 64-neuron network and the
24-neuron 12-channel tree network: V_m, the 7 intracellular and 7 extracellular ions, ATP and the
transmitter levels, each within the stated tolerance of 1e-4 relative (the largest over all
segments is printed per field).

Transmitter levels are not continuous in the trajectory: a release is a draw seeded by
(uint32_t)(V_pre * 1000) (src/simulation_host.cpp:438-440), so a last-bit difference of the
presynaptic potential can move a release by a tick, and the cleft then differs by a whole
release amount. Both runs log every synapse's release ticks; transmitters are compared on the
segments whose incoming synapses released on the same ticks in both, and every other synapse
must keep its number of releases with each release at most one tick away."""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import f64full
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025
TOL = 1e-4
TICKS = 1000


def _check(net, stims, gstims, label):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    glog = f64full.ReleaseLog(len(net.synapses), DT)
    last = None
    for t in range(TICKS):
        g = gstims(t)
        if g is not last:
            sim.upload_stimuli(g)
            last = g
        sim.run_tick(DT)
        glog(t, sim.bion_data().synapses)
    gseg = sim.bion_data().segments.copy()
    gsp = sim.spikes(len(net.segments))
    sim.close()
    ref = net.copy()
    flog = f64full.ReleaseLog(len(net.synapses), DT)
    fsp = f64full.run(ref, TICKS, DT, stims, after=lambda t, n: flog(t, n.synapses))
    same = glog.same_segments(flog, net.synapses, len(net.segments))
    err = f64full.field_errors(gseg, ref.segments, nt_rows=same)
    moved = [k for k in range(len(net.synapses)) if glog.ticks[k] != flog.ticks[k]]
    print(f"{label}: spikes {int(gsp.sum())} (FP64 {int(fsp.sum())}), releases "
          f"{sum(map(len, glog.ticks))} (FP64 {sum(map(len, flog.ticks))}), {len(moved)} of "
          f"{len(net.synapses)} synapses released on other ticks ({int((~same).sum())} segments' "
          f"transmitters left out)", flush=True)
    for k, (mx, md) in err.items():
        print(f"  {k:14s} max {mx:.2e}  median {md:.2e}", flush=True)
    bad = {k: v[0] for k, v in err.items() if v[0] > TOL}
    assert not bad, bad
    assert abs(int(gsp.sum()) - int(fsp.sum())) <= max(2, gsp.sum() // 100)
    assert len(moved) <= max(2, len(net.synapses) // 20)
    for k in moved:
        a, b = glog.ticks[k], flog.ticks[k]
        assert len(a) == len(b) and max(abs(x - y) for x, y in zip(a, b)) <= 1, (k, a, b)


def test_config1_fields_against_fp64():
    net = bion.config1()
    clamp = H.soma_clamp(net, 200.0)
    g = backend.make_stimuli(current=clamp)
    _check(net, lambda t: H.Stimuli(current=clamp), lambda t: g, "64 neurons x 8 segments")


def test_config2_small_fields_against_fp64():
    net = bion.config2(n_neurons=24, syn_per_neuron=100.0)
    pulses = bion.poisson_soma_pulses(net, 200.0, TICKS, DT, 300.0, 5)
    clamp = H.soma_clamp(net, 150.0)
    cur = [np.concatenate([clamp, p]) for p in pulses]
    gst = [backend.make_stimuli(current=c) for c in cur]
    _check(net, lambda t: H.Stimuli(current=cur[t]), lambda t: gst[t], "24 neurons x 32 segments, 12 channels")
