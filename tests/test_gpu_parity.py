"""HIP tick vs the oracle on identical inputs: bit-exact on every byte of the segment and
synapse records, on spike counts and on the delay ring (integer, index and floating-point
state alike - the device reproduces the reference's arithmetic operation by operation)."""
import ctypes as C

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


@pytest.mark.parametrize("generic", [False, True])
def test_post_walk_paths(generic):
    """The post-side walk with wavefronts of uneven chain lengths: transmitters other than the
    receptors' own in the clefts (summed after the walk), ions 4-6 below the floor, myelinated
    post segments (receptors skipped), and - with `generic` - one post slot holding a channel
    instead of a receptor, which selects the full-dispatch kernel."""
    net = bion.config2(n_neurons=12, syn_per_neuron=70.0, seed=31)
    rng = np.random.default_rng(31)
    y = net.synapses
    pick = rng.random(len(y)) < 0.3
    for k in (2, 3, 5, 9, 14):
        sel = pick & (rng.random(len(y)) < 0.5)
        y["cleft_neurotransmitter"][sel, k] = rng.uniform(0.01, 0.5, sel.sum()).astype(np.float32)
    s = net.segments
    low = rng.random(len(s)) < 0.2
    for k in (4, 5, 6):
        s["intra_ions"][low, k] = 1e-12
    posts = np.unique(y["post_segment"])
    for i in rng.choice(posts, 6, replace=False):   # myelin modifier (type 4) on a free slot
        c = s["modifier_count"][i]
        if c < s["modifiers"].shape[1]:
            s["modifiers"][i, c]["type"] = 4
            s["modifier_count"][i] = c + 1
    if generic:
        j = int(np.flatnonzero(y["post_receptor_count"] == 1)[0])
        y["post_receptors"][j, 1]["type"] = bion.K2P
        y["post_receptors"][j, 1]["density"] = 4.17
        y["post_receptor_count"][j] = 2
    pulses = bion.poisson_soma_pulses(net, 200.0, 120, DT, 300.0, 32)
    clamp = H.soma_clamp(net, 150.0)
    cur = [np.concatenate([clamp, p]) for p in pulses]
    gst = [backend.make_stimuli(current=c) for c in cur]
    _compare(net, 120, lambda t: H.Stimuli(current=cur[t]), lambda t: gst[t])


@pytest.mark.parametrize("with_syn", [False, True])
def test_transmitter_liveness(with_syn):
    """Transmitter diffusion where it is live: without synapses the extracellular levels diffuse
    (a ligand injection seeds them); intra_nt becomes live with an intracellular injection
    after some ticks (until then it is +0.0 everywhere and its diffusion is skipped)."""
    net = bion.config1(seed=12) if with_syn else bion.config3(n_cables=4, seg_per_cable=24, seed=13)
    ex = np.zeros(1, bion.LIG_INJ_DT)
    ex["segment_id"], ex["nt_idx"], ex["amount"] = 3, 2, 0.7
    inj = np.zeros(1, bion.LIG_INJ_DT)
    inj["segment_id"], inj["nt_idx"], inj["amount"], inj["intracellular"] = 5, 4, 0.9, 1

    def kinds(t):
        return dict(lig=ex if t == 4 else (inj if t == 15 else None))
    gst = [backend.make_stimuli(**kinds(t)) for t in range(50)]
    _compare(net, 50, lambda t: H.Stimuli(**kinds(t)), lambda t: gst[t])


def test_all_tile_classes_and_stimulus_kinds():
    """64-, 128- and 256-lane tiles in one network (8-, 100- and 200-segment islands), with
    current pulses, ramps, a voltage clamp and one-shot ion/ligand injections."""
    a = bion.build_network(6, 8, "chain", [bion.NAV12, bion.KVDR, bion.K2P, bion.NALCN, bion.PUMP],
                           [bion.CAV21], 4.0, "ampa", 5, 21)
    b = bion.config3(n_cables=2, seg_per_cable=100, seed=22)
    c = bion.config5_cell(23, n_seg=200)
    net = _concat([a, b, c])
    cur = H.soma_clamp(net, 220.0)
    ramp = np.zeros(1, bion.RAMP_DT)
    ramp["segment_id"], ramp["start_pA"], ramp["end_pA"], ramp["start_tick"], ramp["end_tick"] = 60, 0.0, 90.0, 5, 120
    vc = np.zeros(1, bion.VCLAMP_DT)
    vc["segment_id"], vc["target_mV"], vc["start_tick"], vc["end_tick"] = 150, -20.0, 30, 60
    ion = np.zeros(1, bion.ION_INJ_DT)
    ion["segment_id"], ion["ion_idx"], ion["amount_mM"] = 70, 2, 0.004
    lig = np.zeros(1, bion.LIG_INJ_DT)
    lig["segment_id"], lig["nt_idx"], lig["amount"] = 12, 0, 0.5

    def kinds(t):
        return dict(current=cur, ramp=ramp, vclamp=vc, ion=ion if t == 20 else None, lig=lig if t == 40 else None)
    gst = [backend.make_stimuli(**kinds(t)) for t in range(150)]
    _compare(net, 150, lambda t: H.Stimuli(**kinds(t)), lambda t: gst[t])


def _concat(nets):
    """Disjoint union of networks (segment, connection and synapse indices shifted)."""
    segs, conns, syns, neurons = [], [], [], []
    off = 0
    sid = 0
    for n in nets:
        s = n.segments.copy()
        s["id"] = np.arange(off, off + len(s))
        c = n.connections.copy()
        c["segment_a"] += off
        c["segment_b"] += off
        y = n.synapses.copy()
        y["pre_segment"] += off
        y["post_segment"] += off
        y["id"] = np.arange(sid + 1, sid + 1 + len(y))
        sid += len(y)
        nr = n.neurons.copy()
        nr["segment_start"] += off
        nr["soma_index"] += off
        segs.append(s); conns.append(c); syns.append(y); neurons.append(nr)
        off += len(s)
    return bion.Network(np.concatenate(neurons), np.concatenate(segs), np.concatenate(conns), np.concatenate(syns))


@pytest.mark.parametrize("seed", [5, 6])
def test_gap_junctions_and_connexins(seed):
    """Electrical synapses (filtered and unfiltered soma-soma junctions, dendritic bridges)
    interleaved in index order with AMPA synapses, Cx36/Cx43 connexins on somas: every byte
    equals the oracle's (which equals the compiled reference) over 300 ticks."""
    net = bion.gap_network(seed)
    cur = H.soma_clamp(net, 150.0)
    cur["amplitude_pA"][::2] = 0.0
    gst = backend.make_stimuli(current=cur)
    _compare(net, 300, lambda t: H.Stimuli(current=cur), lambda t: gst)
    seg, syn, spk, dl, stats = _gpu_run(net.copy(), 1, lambda t: gst)
    # junctions behind synapses that may change either end's Ca ran in the ordered pass
    assert stats.unsupported_synapses == 0 and stats.order_hazard_synapses > 0


def test_compat_coherence_and_host_edits():
    """COMPAT mode: the AoS is current after every tick and host edits between ticks reach the
    device (as the daemon writes V_m into the AoS, src/daemon_server.cpp:1634-1657)."""
    net = bion.config1(seed=9)
    clamp = H.soma_clamp(net, 200.0)
    o = H.OracleSim(net.copy())
    sim = backend.Simulation(coherence=backend.COHERENCE_COMPAT)
    assert sim.load_network(net.copy()) == 0
    sim.upload_stimuli(backend.make_stimuli(current=clamp))
    view = sim.bion_data()
    for t in range(60):
        if t == 30:   # host edit between ticks on both sides
            o.net.segments["V_m"][5] = -20.0
            view.segments["V_m"][5] = -20.0
        o.step(DT, H.Stimuli(current=clamp))
        sim.run_tick(DT)
        assert H.first_mismatch(o.net.segments, view.segments) is None, (t, H.first_mismatch(o.net.segments, view.segments))
    assert H.first_mismatch(o.net.synapses, view.synapses) is None
    sim.close()


def test_load_bion_file_and_error_codes(tmp_path):
    net = bion.config1(seed=4)
    p = str(tmp_path / "n.bion")
    bion.write_bion(p, net)
    sim = backend.Simulation()
    assert sim.load_bion(p) == 0
    assert sim.load_bion(str(tmp_path / "missing.bion")) == -1
    bad = tmp_path / "bad.bion"
    bad.write_bytes(b"\0" * 200)
    assert sim.load_bion(str(bad)) == -2
    raw = bytearray(open(p, "rb").read())
    raw[4:8] = (15).to_bytes(4, "little")
    (tmp_path / "v15.bion").write_bytes(bytes(raw))
    assert sim.load_bion(str(tmp_path / "v15.bion")) == -3
    raw[4:8] = (13).to_bytes(4, "little")   # v13 header whose sizes name v14 records
    (tmp_path / "v13.bion").write_bytes(bytes(raw))
    assert sim.load_bion(str(tmp_path / "v13.bion")) == 0
    out = str(tmp_path / "saved.bion")
    assert sim.load_bion(p) == 0
    sim.run_ticks(DT, 5)
    assert sim.save_bion(out) == 0
    back = bion.read_bion(out)
    np.testing.assert_array_equal(back.segments["V_m"], sim.bion_data().segments["V_m"])
    sim.close()


@pytest.mark.parametrize("glia", [False, True])
def test_order_dependent_step3_bitexact(glia):
    """Immediate releases behind synapses onto their presynaptic soma and gap junctions behind
    Ca-changing synapses: the ordered pass reproduces the reference's synapse-index order
    (hazard_network), byte for byte over 400 ticks."""
    net = bion.hazard_network(3, glia=glia)
    clamp = H.soma_clamp(net, 200.0)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    assert sim.stats().order_hazard_synapses > 0   # gap junctions and delay-0 releases: any dt
    sim.close()
    st = backend.make_stimuli(current=clamp)
    o = _compare(net, 400, lambda t: H.Stimuli(current=clamp), lambda t: st)
    now = (o.net.synapses["delay"] <= 0.01) & (o.net.synapses["type"] == 0)
    assert (o.net.synapses["RRP"][now] < o.net.synapses["RRP_max"][now]).any()


def test_ordered_pass_planned_per_dt():
    """A 0.01 ms delay is immediate at dt = 0.025 ms and delayed at dt = 0.01 ms: the ordered
    pass is planned for the dt of each tick."""
    net = bion.hazard_network(5)
    clamp = H.soma_clamp(net, 200.0)
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    sim.upload_stimuli(backend.make_stimuli(current=clamp))
    sim.run_tick(0.025)
    a = sim.stats()
    sim.run_tick(0.01)
    b = sim.stats()
    sim.close()
    assert a.order_hazard_synapses > b.order_hazard_synapses > 0
    assert a.ordered_events > 0
    o = H.OracleSim(net.copy())
    o.step(0.025, H.Stimuli(current=clamp))
    o.step(0.01, H.Stimuli(current=clamp))
    g = _gpu_two_dt(net.copy(), clamp)
    assert H.first_mismatch(o.net.segments, g[0]) is None
    assert H.first_mismatch(o.net.synapses, g[1]) is None


def _gpu_two_dt(net, clamp):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net) == 0
    sim.upload_stimuli(backend.make_stimuli(current=clamp))
    sim.run_tick(0.025)
    sim.run_tick(0.01)
    out = sim.bion_data()
    res = (out.segments.copy(), out.synapses.copy())
    sim.close()
    return res


def test_registered_plugins_are_refused():
    """A modifier type or release modulator with a registered plugin function runs that function
    on the reference's CPU tick (include/kernel_dispatch.h:337-343, 575-582) and the built-in one
    on the device: a network with such an instance is refused unless allowed, and loads again
    once the registration is cleared."""
    lib = backend.load_library()
    mod_fn = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_void_p, C.c_float)(lambda m, s, dt: 0.0)
    rel_fn = C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_void_p, C.c_float)(lambda m, c, dt: 1.0)
    lib.register_modifier_current.argtypes = [C.c_uint8, C.c_void_p]
    lib.register_modulator_release.argtypes = [C.c_uint8, C.c_void_p]
    net = bion.config1(seed=9)
    n_k2p = int((net.segments["modifiers"]["type"] == bion.K2P).sum())
    try:
        lib.register_modifier_current(bion.K2P, C.cast(mod_fn, C.c_void_p))
        assert lib.bchip_registered_modifier(bion.K2P) == 1
        sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
        assert sim.load_network(net.copy()) == backend.ERR_UNSUPPORTED
        assert sim.stats().unsupported_modifiers == n_k2p
        sim.allow_unsupported(True)
        assert sim.load_network(net.copy()) == 0
        sim.close()
    finally:
        lib.register_modifier_current(bion.K2P, None)
    mnet = bion.modulated_network(1)
    assert (mnet.synapses["pre_autoreceptor_count"] > 0).any()
    t = int(mnet.synapses["pre_autoreceptors"]["type"][mnet.synapses["pre_autoreceptor_count"] > 0][0, 0])
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    try:
        lib.register_modulator_release(t, C.cast(rel_fn, C.c_void_p))
        assert sim.load_network(mnet.copy()) == backend.ERR_UNSUPPORTED
        assert sim.stats().unsupported_modifiers > 0
    finally:
        lib.register_modulator_release(t, None)
    assert sim.load_network(mnet.copy()) == 0
    assert sim.stats().unsupported_modifiers == 0
    sim.close()


def test_unsupported_work_is_refused(tmp_path):
    """A network that needs work the device does not do loads only on request: a modifier type
    outside the device set (through the drop-in file load) and a post receptor of such a type
    (in-memory load) return ERR_UNSUPPORTED, with the counts in the stats, unless
    bchip_allow_unsupported was set."""
    lib = backend.load_library()
    bad = next(t for t in range(1, 212) if not lib.bchip_modifier_supported(t))
    net = bion.config1(seed=9)
    c = net.segments["modifier_count"][3]
    net.segments["modifiers"]["type"][3, c] = bad
    net.segments["modifier_count"][3] = c + 1
    path = str(tmp_path / "unsupported.bion")
    bion.write_bion(path, net)
    h = lib.gpu_backend_create()
    try:
        assert lib.gpu_backend_load_bion(h, path.encode()) == backend.ERR_UNSUPPORTED
        lib.bchip_allow_unsupported(h, 1)
        assert lib.gpu_backend_load_bion(h, path.encode()) == 0
    finally:
        lib.gpu_backend_destroy(h)
    net = bion.config1(seed=9)
    net.synapses["post_receptors"]["type"][0, 0] = bad
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == backend.ERR_UNSUPPORTED
    assert sim.stats().unsupported_modifiers == 1 and sim.stats().unsupported_synapses == 0
    sim.allow_unsupported(True)
    assert sim.load_network(net) == 0
    sim.close()
