"""Stimulus conventions of the reference's callers.

- Convention 1 (the SYCL backend matches stimulus windows one tick later,
  src/simulation.cpp:201,234) equals convention 0 (the CPU tick and the CUDA backend,
  src/simulation_host.cpp:186; src/gpu_backend_cuda.cu:252-255) run with every window moved one
  tick earlier, bit for bit; and it is not the unshifted run.
- The daemon's own stimulus loop (src/daemon_server.cpp:1634-1657: voltage clamps written into
  V_m, pulses and ramps added to I_buffer in amperes, ion and ligand amounts added to the record
  before each GPU tick) reaches the device in COMPAT coherence exactly as the same writes reach
  the oracle's records.
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _stimuli(net, shift=0):
    somas = np.nonzero(net.segments["segment_type"] == 1)[0]
    cur = np.zeros(len(somas), bion.CURRENT_STIM_DT)
    cur["segment_id"] = net.segments["id"][somas]
    cur["amplitude_pA"] = 220.0
    cur["duration_ms"] = 0.5
    cur["start_tick"] = 7 - shift
    ramp = np.zeros(2, bion.RAMP_DT)
    ramp["segment_id"] = net.segments["id"][[1, 9]]
    ramp["start_pA"], ramp["end_pA"] = 0.0, 90.0
    ramp["start_tick"], ramp["end_tick"] = 3 - shift, 40 - shift
    vc = np.zeros(1, bion.VCLAMP_DT)
    vc["segment_id"] = net.segments["id"][17]
    vc["target_mV"] = -30.0
    vc["start_tick"], vc["end_tick"] = 12 - shift, 25 - shift
    return dict(current=cur, ramp=ramp, vclamp=vc)


def _run(net, convention, stim, ticks=60):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.set_stimulus_convention(convention)
    sim.set_stimuli(0, **stim)
    sim.run_ticks(DT, ticks)
    out = sim.bion_data()
    res = out.segments.copy(), out.synapses.copy(), sim.spikes(len(net.segments))
    sim.close()
    return res


def test_sycl_convention_is_one_tick_later():
    net = bion.config1(seed=21)
    a = _run(net, 1, _stimuli(net))
    b = _run(net, 0, _stimuli(net, shift=1))
    c = _run(net, 0, _stimuli(net))
    assert H.first_mismatch(a[0], b[0]) is None, H.first_mismatch(a[0], b[0])
    assert H.first_mismatch(a[1], b[1]) is None
    np.testing.assert_array_equal(a[2], b[2])
    assert H.first_mismatch(a[0], c[0]) is not None
    sim = backend.Simulation()
    with pytest.raises(ValueError):
        sim.set_stimulus_convention(2)
    sim.close()


def test_daemon_writes_into_the_records():
    """The daemon's loop on both sides: writes into the records, then a tick without stimuli."""
    net = bion.config1(seed=22)
    st = _stimuli(net)
    ion = np.zeros(1, bion.ION_INJ_DT)
    ion["segment_id"], ion["ion_idx"], ion["amount_mM"] = net.segments["id"][3], 2, 0.002
    lig = np.zeros(1, bion.LIG_INJ_DT)
    lig["segment_id"], lig["nt_idx"], lig["amount"] = net.segments["id"][4], 0, 0.5
    o = H.OracleSim(net.copy())
    sim = backend.Simulation(coherence=backend.COHERENCE_COMPAT)
    assert sim.load_network(net.copy()) == 0
    view = sim.bion_data()
    ids = net.segments["id"]

    def daemon_writes(seg, t):   # src/daemon_server.cpp:1636-1655, on one record array
        for q in st["vclamp"]:
            if q["start_tick"] <= t < q["end_tick"]:
                seg["V_m"][ids == q["segment_id"]] = q["target_mV"]
        for q in st["current"]:
            dur = np.uint64(np.float32(q["duration_ms"]) / np.float32(DT))
            if q["start_tick"] <= t < q["start_tick"] + dur:
                k = np.nonzero(ids == q["segment_id"])[0]
                seg["I_buffer"][k] += np.float32(q["amplitude_pA"]) * np.float32(1e-12)
        for q in st["ramp"]:
            if q["start_tick"] <= t < q["end_tick"]:
                frac = np.float32(t - q["start_tick"]) / np.float32(q["end_tick"] - q["start_tick"])
                amp = np.float32(q["start_pA"] + (q["end_pA"] - q["start_pA"]) * frac)
                seg["I_buffer"][ids == q["segment_id"]] += amp * np.float32(1e-12)
        if t in (5, 30):
            k = np.nonzero(ids == ion["segment_id"][0])[0]
            seg["intra_ions"][k, 2] += np.float32(ion["amount_mM"][0])
            k = np.nonzero(ids == lig["segment_id"][0])[0]
            seg["neurotransmitter"][k, 0] += np.float32(lig["amount"][0])

    for t in range(50):
        daemon_writes(o.net.segments, t)
        daemon_writes(view.segments, t)
        o.step(DT)
        sim.run_tick(DT)
        assert H.first_mismatch(o.net.segments, view.segments) is None, (t, H.first_mismatch(o.net.segments, view.segments))
    assert H.first_mismatch(o.net.synapses, view.synapses) is None
    np.testing.assert_array_equal(o.spikes(), sim.spikes(len(net.segments)))
    sim.close()
