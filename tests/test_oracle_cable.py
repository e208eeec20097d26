"""The oracle's cable mode and its FP64 restatement (the reference gradient of the adjoint comes
from central differences of the FP64 restatement).

Cable mode holds ion concentrations, ATP and transmitters, sets phi_e = 0 and lets step 2.5 move
V only; each connection's axial conductance and flux carry a scale. It has no reference
counterpart, so it is pinned from two sides:
- on one tick of a single segment with one channel the reference's own tick (the bit-exact
  oracle, normal mode) changes ions only after the channel's current is computed, so the FP64
  restatement must give the reference's V to float precision: this pins the five channel
  restatements against the reference path;
- the oracle in cable mode (the reference's float arithmetic with the ion updates off) and the
  independent FP64 restatement must agree over hundreds of ticks: to float precision for
  channels, and to 1e-4 relative of the voltage change for axial coupling, where the chord term
  s*(inp - G dV) cancels to ~3 digits in the reference's float arithmetic.
"""
import numpy as np

from bioneuroncore_amd import bion
import harness as H

DT = 0.025
TYPES = ["Nav1.2", "Kv-DR", "K2P", "NALCN", "HCN1"]


def _one_segment(types):
    base = bion.config5_cell(62, n_seg=2)
    net = bion.Network(base.neurons, base.segments[:1].copy(), base.connections[:0].copy(), base.synapses)
    m = net.segments["modifiers"]
    keep = m[0][types].copy()
    m[0][:len(types)] = keep
    net.segments["modifier_count"][0] = len(types)
    return net


def _oracle_cable(net, ticks, cur=None, scale=None):
    o = H.OracleSim(net.copy())
    o.set_cable(True, scale)
    trace = []
    for _ in range(ticks):
        o.step(DT, H.Stimuli(current=cur) if cur is not None else None)
        trace.append(o.net.segments["V_m"].astype(np.float64))
    return np.array(trace), o.net.segments


def test_fp64_channels_equal_reference_tick():
    """One tick, one segment, one channel: FP64 restatement == the reference's tick (float V)."""
    for k, name in enumerate(TYPES):
        net = _one_segment([k])
        cur = H.soma_clamp(net, 30.0)
        seg, _, _, _ = H.run_oracle(net, 1, DT, lambda t: H.Stimuli(current=cur))
        L, tr = H.Cable64(net.segments, net.connections, current=cur).loss(DT, 1, [0], np.zeros((1, 1)), trace=True)
        dv = abs(tr[0, 0] - float(seg["V_m"][0]))
        # float V_m rounding (half an ulp of ~65 mV is 3.8e-6) bounds the difference
        assert dv < 4e-6, (name, dv)
        assert abs(tr[0, 0] - float(net.segments["V_m"][0])) > 1e-3, name   # the channel moved V


def test_fp64_channels_match_oracle_cable_mode():
    for types in ([0], [1], [2], [3], [4], [0, 1, 2, 3, 4]):
        net = _one_segment(types)
        cur = H.soma_clamp(net, 30.0)
        T = 200
        tro, seg = _oracle_cable(net, T, cur)
        L, tr = H.Cable64(net.segments, net.connections, current=cur).loss(DT, T, [0], np.zeros((T, 1)), trace=True)
        assert np.abs(tr[:, 0] - tro[:, 0]).max() < 2e-5, types
        # concentrations are held
        np.testing.assert_array_equal(seg["intra_ions"], net.segments["intra_ions"])
        np.testing.assert_array_equal(seg["extra_ions"], net.segments["extra_ions"])


def test_fp64_axial_coupling_matches_oracle_cable_mode():
    rng = np.random.default_rng(1)
    for amp in (10.0, 0.5):
        net = bion.config5_cell(61, n_seg=40)
        net.segments["modifier_count"] = 0
        net.segments["V_m"] = -65 + amp * np.sin(np.arange(len(net.segments)) / 3.0)
        sc = rng.uniform(0.5, 1.5, len(net.connections)).astype(np.float32)
        T = 50
        tro, seg = _oracle_cable(net, T, None, sc)
        L, tr = H.Cable64(net.segments, net.connections, scale=sc).loss(
            DT, T, np.arange(len(net.segments)), np.zeros((T, len(net.segments))), trace=True)
        moved = np.abs(tro[-1] - net.segments["V_m"]).max()
        assert moved > 0.5 * amp
        assert np.abs(tr - tro).max() < 2e-4 * moved + 1e-5, amp


def test_fp64_cell_matches_oracle_cable_mode():
    rng = np.random.default_rng(2)
    net = bion.config5_cell(63, n_seg=200)
    sc = rng.uniform(0.5, 1.5, len(net.connections)).astype(np.float32)
    cur = H.soma_clamp(net, 50.0)
    T = 200
    tro, seg = _oracle_cable(net, T, cur, sc)
    L, tr = H.Cable64(net.segments, net.connections, scale=sc, current=cur).loss(
        DT, T, np.arange(200), np.zeros((T, 200)), trace=True)
    assert np.abs(tr - tro).max() < 5e-4
    np.testing.assert_array_equal(seg["intra_ions"], net.segments["intra_ions"])
    np.testing.assert_array_equal(seg["ATP_level"], net.segments["ATP_level"])


def test_unit_scale_is_the_reference_tick():
    """Normal mode with every scale 1 is bit-identical to the unscaled (reference) tick."""
    net = bion.config5_cell(64, n_seg=60)
    cur = H.soma_clamp(net, 80.0)
    a = H.OracleSim(net.copy())
    b = H.OracleSim(net.copy())
    b.set_cable(False, np.ones(len(net.connections), np.float32))
    for _ in range(30):
        a.step(DT, H.Stimuli(current=cur))
        b.step(DT, H.Stimuli(current=cur))
    assert H.first_mismatch(a.net.segments, b.net.segments) is None


def test_fp64_loss_definition():
    net = bion.config5_cell(65, n_seg=50)
    cur = H.soma_clamp(net, 40.0)
    T = 20
    tgt = np.linspace(-70, -40, T, dtype=np.float32)[:, None]
    L, tr = H.Cable64(net.segments, net.connections, current=cur).loss(DT, T, [0], tgt, trace=True)
    assert abs(L - np.mean((tr - tgt.astype(np.float64)) ** 2)) < 1e-9 * L
