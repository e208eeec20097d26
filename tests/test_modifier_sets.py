"""Extended modifier sets on the device against the oracle, bit for bit (the oracle equals the
compiled reference on the same networks, tests/test_oracle_vs_reference.py):

- ion homeostasis: the alpha3 and glial Na/K pumps, PMCA, NCX, KCC2, NKCC1, the Ca2+ buffer and
  HCN2 (src/hh_biological.cpp:99-141, 2574-2865, 3414-3445, 3599);
- the channel set of the spinal cell types: developmental Nav1.2, Kv2.1, Kv4.2, Cav1.3, the two
  persistent Na+ channels, Kir2.1 and the ClC-2 leak;
- the ER Ca2+ store (SERCA, IP3R, RyR) and PLC (src/hh_biological.cpp:3473-3591,
  include/modifier_helpers.h:1553-1580);
- the metabotropic G protein, the release site, somatic ATP synthesis and thermoregulation,
  CREB and MRGPR (src/hh_biological.cpp, include/modifier_helpers.h mrgpr_signal);
- the MAPK/ERK, NF-kB, mTORC1, protein-synthesis, cofilin and MMP-9 cascades, the TRPV1
  transmitter receptor and rod phototransduction (src/hh_biological.cpp);
- CaMKII, calcineurin, PP1 / DARPP-32 and PKC, whose activities step 1 and step 3 read as they
  stand (src/hh_biological.cpp:545-591,3222-3253, include/modifier_helpers.h:1584-1593,
  include/kernel_dispatch.h:612-631), and the IDs without code in the reference;
- the physical, hunger, light, olfactory, vibration, otolith, cupula and cryptochrome sensors and
  thermal noise (src/hh_biological.cpp:483-540, 2181-2440, 3122-3141);
- a network that mixes them with the core channel set runs the extended step-1 kernel, and the
  core-only networks keep the lean one (results identical either way).
"""
import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _compare(net, ticks, cur):
    o = H.OracleSim(net.copy())
    for _ in range(ticks):
        o.step(DT, H.Stimuli(current=cur))
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.upload_stimuli(backend.make_stimuli(current=cur))
    sim.run_ticks(DT, ticks)
    out = sim.bion_data()
    seg, syn, spk, stats = out.segments.copy(), out.synapses.copy(), sim.spikes(len(net.segments)), sim.stats()
    sim.close()
    print(f"unsupported={stats.unsupported_modifiers} spikes={int(spk.sum())}", flush=True)
    assert stats.unsupported_modifiers == 0
    assert H.first_mismatch(o.net.segments, seg) is None, H.first_mismatch(o.net.segments, seg)
    assert H.first_mismatch(o.net.synapses, syn) is None, H.first_mismatch(o.net.synapses, syn)
    np.testing.assert_array_equal(o.spikes(), spk)
    return seg, spk


@pytest.mark.parametrize("seed", [31, 32])
def test_ion_homeostasis_modifiers(seed):
    net = bion.homeostasis_network(seed)
    cur = H.soma_clamp(net, 170.0)
    cur = cur[cur["segment_id"] < 12 * 6]
    _, spk = _compare(net, 300, cur)
    assert spk.sum() > 0


def test_pmca_hill_other_than_two_is_counted():
    net = bion.homeostasis_network(33)
    mt = net.segments["modifiers"]["type"]
    i, j = np.nonzero(mt == bion.PMCA)
    net.segments["modifiers"]["state"][i[0], j[0], 2] = 3.0
    sim = backend.Simulation()
    assert sim.load_network(net.copy()) == backend.ERR_UNSUPPORTED   # refused by default
    assert sim.stats().unsupported_modifiers == 1
    sim.allow_unsupported(True)
    assert sim.load_network(net) == 0
    assert sim.stats().unsupported_modifiers == 1
    sim.close()


@pytest.mark.parametrize("seed", [61, 62])
def test_er_store_and_plc(seed):
    """SERCA, IP3R and RyR with the segment's ER Ca2+ store, PLC fed by mGluR5 (src/hh_biological.cpp:
    3473-3591, include/modifier_helpers.h:1553-1580), 400 ticks: every byte, ER_Ca2 included."""
    net = bion.er_network(seed)
    cur = H.soma_clamp(net, 160.0)
    seg, spk = _compare(net, 400, cur)
    assert spk.sum() > 0
    er0 = net.segments["ER_Ca2"]
    assert np.any(seg["ER_Ca2"] > er0) and np.any(seg["ER_Ca2"] < er0)


@pytest.mark.parametrize("seed", [71, 72])
def test_signalling_modifiers(seed):
    """The metabotropic G protein (3), the release site (5), somatic ATP synthesis (77) and
    thermoregulation (78), CREB (79) and MRGPR (188) feeding PLC, 400 ticks: every byte of the
    records, temperature and transcription signal included."""
    net = bion.signalling_network(seed)
    cur = H.soma_clamp(net, 160.0)
    seg, spk = _compare(net, 400, cur)
    assert spk.sum() > 0
    assert np.any(seg["transcription_signal"] > 0.0)


@pytest.mark.parametrize("seed", [81, 82])
def test_cascade_modifiers(seed):
    """MAPK/ERK (raising CREB), NF-kB, mTORC1, protein synthesis, cofilin and MMP-9, the TRPV1
    transmitter receptor and rod phototransduction, 400 ticks: every byte of the records."""
    net = bion.cascade_network(seed)
    cur = H.soma_clamp(net, 160.0)
    seg, spk = _compare(net, 400, cur)
    assert spk.sum() > 0


@pytest.mark.parametrize("seed", [91, 92])
def test_sensor_modifiers(seed):
    """The sensors (6, 94, 95, 108, 115-117), thermal noise seeded by Segment.id and V (111) and
    the cryptochrome's boost of the first light sensor (120, before and after it in slot order),
    400 ticks: every byte of the records."""
    net = bion.sensor_network(seed)
    cur = H.soma_clamp(net, 160.0)
    seg, spk = _compare(net, 400, cur)
    assert spk.sum() > 0


@pytest.mark.parametrize("seed", [91, 92])
def test_kinase_modifiers(seed):
    """CaMKII, calcineurin, PP1 / DARPP-32 and PKC read live by Nav1.2, Cx36, Kv4.2, vesicle
    recovery and plasticity, with the IDs that have no code in the reference, 400 ticks: every
    byte of the records."""
    net = bion.kinase_network(seed)
    cur = H.soma_clamp(net, 160.0)
    seg, spk = _compare(net, 400, cur)
    assert spk.sum() > 0


def test_msn_receptors_read_the_d1_d2_modulators():
    """The MSN NMDA/AMPA receptors (75/76) scale their current by state[1] of the segment's first
    D1 (65) / D2 (66) modulator, floored at 0 (receptors_biological.cpp:337-369). The modulators
    themselves need glibc pow and are outside the device set, so the network loads only with
    bchip_allow_unsupported and both the device and the oracle hold their states; the device's
    scaling must then equal the oracle's byte for byte, 200 ticks."""
    net = bion.kinase_network(93)
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(93)
    somas = np.nonzero(segs["segment_type"] == 1)[0]
    for q, i in enumerate(somas):
        k = int(segs["modifier_count"][i])
        for j, t in enumerate((65, 66)):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = 1.0
            mods["state"][i, k + j, 1] = -0.3 if q % 4 == 0 else rng.uniform(0.2, 1.8)
        segs["modifier_count"][i] = k + 2
    cur = H.soma_clamp(net, 160.0)
    o = H.OracleSim(net.copy())
    for _ in range(200):
        o.step(DT, H.Stimuli(current=cur))
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == backend.ERR_UNSUPPORTED
    sim.allow_unsupported(True)
    assert sim.load_network(net.copy()) == 0
    assert sim.stats().unsupported_modifiers == 2 * len(somas)
    sim.upload_stimuli(backend.make_stimuli(current=cur))
    sim.run_ticks(DT, 200)
    out = sim.bion_data()
    seg, syn = out.segments.copy(), out.synapses.copy()
    sim.close()
    assert H.first_mismatch(o.net.segments, seg) is None, H.first_mismatch(o.net.segments, seg)
    assert H.first_mismatch(o.net.synapses, syn) is None, H.first_mismatch(o.net.synapses, syn)


@pytest.mark.parametrize("seed,myelin", [(41, False), (42, False), (43, True)])
def test_channel_set_of_the_spinal_cell_types(seed, myelin):
    """Developmental Nav1.2, Kv2.1, Kv4.2, Cav1.3, both persistent Na+ channels, Kir2.1, ClC-2
    (src/hh_biological.cpp:142-185, 314-414, 789-836, 1392-1428, 2951-3037); myelin blocks the
    listed types (include/kernel_dispatch.h:299-309)."""
    net = bion.channel_network(seed, myelin=myelin)
    cur = H.soma_clamp(net, 120.0)
    _, spk = _compare(net, 400, cur)
    assert spk.sum() > 0


@pytest.mark.parametrize("seed", [51, 52])
def test_receptor_families(seed):
    """Glycine, nicotinic, 5-HT3, kainate, alpha7 and GluN2B receptors on synapses and segments;
    Gi/o- and Gs-coupled receptors as segment and synaptic modifiers
    (src/receptors_biological.cpp:40-200, include/modifier_helpers.h:1332-1398)."""
    net = bion.receptor_network(seed)
    cur = H.soma_clamp(net, 150.0)
    _, spk = _compare(net, 400, cur)
    assert spk.sum() > 0


def test_remaining_gpcr_families():
    """bion.GPCR_REST as segment modifiers reading injected transmitters (gpcr_network)."""
    net = bion.gpcr_network(4)
    clamp = H.soma_clamp(net, 120.0)
    o = H.OracleSim(net.copy())
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    assert sim.stats().unsupported_modifiers == 0
    for t in range(240):
        lig = bion.gpcr_ligands(net, t)
        o.step(DT, H.Stimuli(current=clamp, lig=lig))
        sim.upload_stimuli(backend.make_stimuli(current=clamp, lig=lig))
        sim.run_tick(DT)
    g = sim.bion_data()
    assert H.first_mismatch(o.net.segments, g.segments) is None, H.first_mismatch(o.net.segments, g.segments)
    np.testing.assert_array_equal(o.spikes(), sim.spikes(len(net.segments)))
    sim.close()
