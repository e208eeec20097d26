"""The oracle against the committed golden fixtures (written by tests/golden/gen_golden.py from
the reference's own compiled CPU library). Runs anywhere - no reference checkout needed."""
import hashlib
import os

import numpy as np
import pytest

from bioneuroncore_amd import bion
import harness as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DT = 0.025
SEG_FIELDS = ["V_m", "intra_ions", "extra_ions", "ATP_level", "neurotransmitter", "radius", "length", "segment_type"]


def _kat():
    return np.load(os.path.join(GOLD, "modifier_kat.npz"))


@pytest.mark.parametrize("mtype", [7, 8, 9, 11, 103, 104, 107, 112, 123, 129, 133, 140, 141, 142, 150, 156])
def test_modifier_known_answers(mtype):
    """Each covered channel/receptor, single call, bit-exact: return value, all 8 states, ions, ATP, V_m."""
    g = _kat()
    lib = H.oracle_lib()
    mods = g[f"t{mtype}_mod_in"].copy()
    dts = g[f"t{mtype}_dt"]
    n = len(mods)
    segs = np.zeros(n, bion.SEGMENT_DT)
    for f in SEG_FIELDS:
        segs[f] = g[f"t{mtype}_seg_in.{f}"]
    ret = np.zeros(n)
    for i in range(n):
        ret[i] = lib.oracle_modifier_current(mtype, mods[i:i + 1].ctypes.data, segs[i:i + 1].ctypes.data, float(dts[i]))
    assert ret.tobytes() == g[f"t{mtype}_ret"].tobytes()
    assert mods.tobytes() == g[f"t{mtype}_mod_out"].tobytes()
    for f in SEG_FIELDS:
        assert segs[f].tobytes() == g[f"t{mtype}_seg_out.{f}"].tobytes(), f


def _replay(path, net, stim_fn, ticks):
    g = np.load(path)
    assert str(g["net_hash"]) == hashlib.sha256(net.segments.tobytes() + net.synapses.tobytes()).hexdigest(), \
        "network builder drifted from the fixture"
    sim = H.OracleSim(net.copy())
    prev = np.zeros(len(net.segments), np.uint32)
    spikes, rels, vs, hs, hy = [], [], [], [], []
    somas = np.nonzero(net.segments["segment_type"] == 1)[0]
    for t in range(ticks):
        sim.step(DT, stim_fn(t))
        sp = sim.spikes()
        spikes += [(t, int(i)) for i in np.nonzero(sp != prev)[0]]
        prev = sp
        rels += [(t, int(y)) for y in np.nonzero(sim.net.synapses["refractory"] == 31)[0]]
        if t % 10 == 9:
            vs.append(sim.net.segments["V_m"][somas].copy())
        if t % 100 == 99:
            hs.append(hashlib.sha256(sim.net.segments.tobytes()).hexdigest())
            hy.append(hashlib.sha256(sim.net.synapses.tobytes()).hexdigest())
    np.testing.assert_array_equal(np.array(spikes, np.int64).reshape(-1, 2), g["spike_events"])
    np.testing.assert_array_equal(np.array(rels, np.int64).reshape(-1, 2), g["release_events"])
    assert np.array(vs).tobytes() == g["v_soma"].tobytes()
    assert list(hs) == list(g["seg_hash"]) and list(hy) == list(g["syn_hash"])
    return g


def test_trace_config1_4000_ticks():
    net = bion.config1()
    clamp = H.soma_clamp(net, 200.0)
    g = _replay(os.path.join(GOLD, "trace_config1.npz"), net, lambda t: H.Stimuli(current=clamp), 4000)
    assert len(g["spike_events"]) > 1000 and len(g["release_events"]) > 100


def test_trace_config2_small_1000_ticks():
    net = bion.config2(n_neurons=24, syn_per_neuron=100.0)
    pulses = bion.poisson_soma_pulses(net, 200.0, 1000, DT, 300.0, 5)
    clamp = H.soma_clamp(net, 150.0)
    g = _replay(os.path.join(GOLD, "trace_config2s.npz"), net,
                lambda t: H.Stimuli(current=np.concatenate([clamp, pulses[t]])), 1000)
    assert len(g["release_events"]) > 100
