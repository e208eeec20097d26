"""Throughput of the per-tick step on one MI355X.

Workload (default): 100k neurons x (soma + 31-node binary dendrite tree) = 3.2M segments,
12 modifiers per segment (log-normal densities), 10M chemical synapses (fan-in 100 from
uniform presynaptic somas, 70% AMPA+NMDA / 30% GABA-A, 1-20 tick delays), somatic Poisson
current pulses at 5 Hz, dt = 0.025 ms, FP64/FP32 mixed state as the reference's ABI.
A step is one tick: upload of that tick's stimuli + gpu_backend_run_tick (synchronous).

Prints one JSON line. value = segment-ticks/s over the timed steps (inputs resident in HBM).
roofline: the dominant kernel's algorithmic bytes per launch / its mean HIP-event duration.
cpu_baseline: the reference CPU tick (oracle/_ref, compiled from the reference sources) or,
when absent, the oracle restatement, on a bounded sample, 1 thread.
"""
import argparse
import ctypes as C
import json
import os
import resource
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bioneuroncore_amd import backend, bion  # noqa: E402

HBM_PEAK_GBS = 8000.0
PHASES = ("k_modifiers", "k_diffusion", "k_syn_pre", "k_syn_post", "k_syn_plast", "k_syn_rrp")


def algorithmic_bytes(net, nt_ex_live, nt_in_live):
    """Compulsory HBM bytes per launch of each kernel: every dynamic field it owns read once and
    written once, every static parameter it needs read once, at the ABI's precision (f64 V_m,
    ions and modifier states; f32 elsewhere). Transmitter fields count only where the kernel
    must move them (the electrodiffusion kernel skips them when they are dead)."""
    n = len(net.segments)
    nc = len(net.connections)
    ns = len(net.synapses)
    types = net.segments["modifiers"]["type"]
    count = net.segments["modifier_count"]
    # modifier states per type: gates read and written, parameters only read
    gate_w = {142: 2, 141: 2, 156: 2, 150: 3, 112: 1, 140: 2, 129: 1, 103: 1, 104: 1, 107: 0, 7: 0, 133: 0, 123: 1}
    param_r = {142: 1, 141: 1, 156: 2, 150: 1, 112: 5, 140: 0, 129: 5, 103: 3, 104: 2, 107: 2, 7: 0, 133: 1, 123: 3}
    gates = params = 0
    for t in np.unique(types[types > 0]):
        c = int((types == t).sum())
        gates += c * gate_w.get(int(t), 0)
        params += c * param_r.get(int(t), 0)
    nmod = int(count.sum())
    rec = int(net.synapses["post_receptor_count"].astype(np.int64).sum())
    ion = 8 + 7 * 8 + 7 * 8          # V_m + 7 intra + 7 extra (f64)
    nt_planes = 15 * 4 * 2           # 15 f32 transmitters, read + written
    return {
        # R: V, ions, ATP, area, vol, PKC/PKA/cAMP, flags, count; W: V, ions, ATP, I_buffer;
        # per modifier type, active-vars and density; gate states R+W, parameter states R
        "k_modifiers": n * (ion + 4 + 4 + 8 + 12 + 2) + n * (ion + 4 + 8)
        + nmod * (1 + 1 + 8) + gates * 16 + params * 8,
        # R: V, ions, I_buffer, area, cap, vol, radius, length, CSR offset, ATP/temp/txn;
        # W: V, ions, ATP/temp/txn; per connection its end pair, intracellular flag and two
        # incidence entries
        "k_diffusion": n * (ion + 8 + 4 + 4 + 8 + 4 + 4 + 4 + 12) + n * (ion + 12)
        + (n * nt_planes if nt_ex_live else 0) + (n * nt_planes if nt_in_live else 0) + nc * (4 + 1 + 8),
        # per synapse: flags, refractory R+W, delay mask, pre index, summary index, bouton Ca
        # R+W, P_rel, pre area, id, RRP, RRP_max, delay, amount, receptor count/type, release
        # probability out; per presynaptic segment its 64-byte summary
        "k_syn_pre": ns * (1 + 2 + 1 + 4 + 4 + 8 + 4 + 4 + 4 + 4 + 4 + 4 + 4 + 2 + 4),
        # per segment R: V, ions 0-6 intra, 0-3 extra, I_buffer, area, vol, flags, cap, lane map,
        # chain length; W: ions 0-3, 15 transmitters, V. Per synapse: flags, mask, one live
        # cleft transmitter, receptor count, post area, current out, second-walk flags + mask;
        # per receptor: type, active-vars, density, two states R+W
        "k_syn_post": n * (8 + 7 * 8 + 4 * 8 + 8 + 4 + 8 + 1 + 4 + 4 + 4) + n * (4 * 8 + 15 * 4 + 8)
        + ns * (1 + 2 + 4 + 1 + 4 + 4 + 3) + rec * (1 + 1 + 8 + 2 * 2 * 8),
        # flags, current, mask, tau, pp1 read; spine Ca, one cleft transmitter, BCM, CaMKII,
        # calcineurin, weight read + written; depression written
        "k_syn_plast": ns * (1 + 4 + 2 + 4 + 4 + 6 * 2 * 4 + 4),
        # flags, release probability, RRP R+W, RRP_max, source index, recovery rate
        "k_syn_rrp": ns * (1 + 4 + 2 * 4 + 4 + 4 + 4),
    }


def tick_bytes_model(net):
    """Whole-tick compulsory bytes at the ABI's precision: 621 B per segment (504 dynamic R+W,
    117 static), 12 B per connection, 193 B per synapse (the configuration-2 unit figures)."""
    return len(net.segments) * 621 + len(net.connections) * 12 + len(net.synapses) * 193


def measured_traffic(kernel, neurons):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    of this workload (already scaled by the copy-kernel calibration), or None when the summary
    does not cover this kernel or size."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_tick_config2_3p2M.md")
    if neurons != 100_000 or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a SIMD for 4 cycles (FP32 and FP64
# alike without packed math); 2.4 GHz peak engine clock.
VALU_SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def measured_valu(kernel, neurons):
    """VALU instructions per launch of `kernel` (wavefront instructions, SQ_INSTS_VALU) from the
    committed rocprofv3 counter summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "sq_counters_tick_config2_3p2M.md")
    if neurons != 100_000 or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == kernel:
            try:
                return float(cells[3])
            except ValueError:
                return None
    return None


def build_workload(neurons, ticks, seed=0xB10C0002):
    net = bion.config2(n_neurons=neurons, seed=seed)
    pulses = bion.poisson_soma_pulses(net, 5.0, ticks, 0.025, 1000.0, seed + 1)
    stims = [backend.make_stimuli(current=p[:256]) for p in pulses]
    return net, stims


def _cpu_stepper(net, stim_fn, kind):
    """step(t): one CPU tick of `net` in place with stim_fn(t) (harness Stimuli or None) -
    the reference (oracle/_ref, compiled from the reference's sources) or the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    P = C.c_void_p
    if kind == "reference":
        lib = C.CDLL(H.REF_SO)
        lib.core_run_tick.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, C.c_uint8, P, C.c_float] + [P] * 12
        tick = np.zeros(1, np.uint64)
        mech = np.zeros(4, np.uint8)
        mc = np.zeros(1, np.uint32)

        def step(t):
            st = stim_fn(t) if stim_fn else None
            args = st.args() if st is not None else [None] * 10
            lib.core_run_tick(net.segments.ctypes.data, len(net.segments), net.connections.ctypes.data,
                              len(net.connections), net.synapses.ctypes.data, len(net.synapses), 0,
                              tick.ctypes.data, 0.025, *args, mech.ctypes.data, mc.ctypes.data)
        return step
    sim = H.OracleSim(net)
    return lambda t: sim.step(0.025, stim_fn(t) if stim_fn else None)


def _cpu_kind():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    return "reference" if H.ref_available() else "port"


def _gpu_stims_as_cpu(stims):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    return lambda t: H.Stimuli(current=stims[t]["current_stim"][0][:stims[t]["current_stim_count"][0]])


def cpu_baseline(budget_s=12.0, neurons=250):
    """Reference CPU tick on a bounded sample of the same workload (250 neurons = 8000 segments,
    25k synapses), one thread, ticks until the time budget is used."""
    net, stims = build_workload(neurons, 4000, seed=0xB10C0102)
    kind = _cpu_kind()
    step = _cpu_stepper(net, _gpu_stims_as_cpu(stims), kind)
    step(0)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget_s and k < 3999:
        step(k + 1)
        k += 1
    el = time.perf_counter() - t0
    return {"value": len(net.segments) * k / el, "unit": "segment-ticks/s", "cores": 1, "kind": kind,
            "sample": f"{neurons} neurons x 32 segments ({len(net.segments)} seg, {len(net.synapses)} syn), "
                      f"{k} ticks in {el:.1f}s, same builder and stimuli as the GPU workload"}


def cpu_full(neurons, ticks, warmup=2):
    """The CPU tick on the full workload (default 100k neurons = 3.2M segments, 10M synapses),
    one thread: `warmup` untimed ticks, then `ticks` timed ones, each printed as it ends."""
    t0 = time.time()
    net, stims = build_workload(neurons, warmup + ticks)
    kind = _cpu_kind()
    print(f"# cpu full size: built {len(net.segments)} segments in {time.time() - t0:.1f}s ({kind})",
          file=sys.stderr, flush=True)
    step = _cpu_stepper(net, _gpu_stims_as_cpu(stims), kind)
    times = []
    for t in range(warmup + ticks):
        t1 = time.perf_counter()
        step(t)
        dt_ = time.perf_counter() - t1
        if t >= warmup:
            times.append(dt_)
        print(f"# cpu tick {t}: {dt_:.2f}s", file=sys.stderr, flush=True)
    return {"kind": kind, "cores": 1, "segments": len(net.segments), "synapses": len(net.synapses),
            "warmup": warmup, "ticks": ticks, "s_per_tick": float(np.mean(times)),
            "segment_ticks_per_s": len(net.segments) / float(np.mean(times))}


def cpu_calibration():
    """t(oracle) / t(reference) for the same ticks of the same networks, one thread each: the
    64-neuron chain network (4000 ticks, 200 pA soma clamp) and the 12-channel tree network at
    1/100 scale (1000 neurons, 32k segments, 40 ticks of its Poisson pulses)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    if not H.ref_available():
        return {"error": "oracle/_ref not built"}
    out = {}
    c1 = bion.config1()
    clamp = H.soma_clamp(c1, 200.0)
    n2, st2 = build_workload(1000, 40)
    for name, net, fn, ticks in (("config1_4000_ticks", c1, lambda t: H.Stimuli(current=clamp), 4000),
                                 ("config2_1000_neurons_40_ticks", n2, _gpu_stims_as_cpu(st2), 40)):
        el = {}
        for kind in ("reference", "port"):
            x = net.copy()
            step = _cpu_stepper(x, fn, kind)
            t0 = time.perf_counter()
            for t in range(ticks):
                step(t)
            el[kind] = time.perf_counter() - t0
            print(f"# calibration {name} {kind}: {el[kind]:.2f}s", file=sys.stderr, flush=True)
        out[name] = {"reference_s": el["reference"], "oracle_s": el["port"],
                     "oracle_over_reference": el["port"] / el["reference"]}
    return out


def _peak_rss_gb():
    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6   # kB -> GB


def batched_vs_plain(n_networks, ticks, dt=0.025, seed=0xB10C0004, graphs=True, plain_ticks=None):
    """The batched sweep: n_networks networks built like the 64-neuron chain network, segment
    counts log-uniform in 2k-200k, each with its own seed and g_max scaling and a 200 pA soma
    clamp. Plain: one simulation per network, ticked one network at a time (each tick is that
    network's own launch sequence and synchronisation), for plain_ticks ticks (a rate: every
    network's tick costs the same at any tick count). Batched: all networks in one simulation,
    autotuned schedule, `ticks` ticks replayed from captured hipGraphs. Both are bit-exact to the
    per-network CPU tick; throughput counts segment-ticks of all networks. Each network is built,
    loaded and dropped in turn for the plain leg and built again for the batch, so the host holds
    at most the networks plus one copy of their records."""
    neurons = bion.config4_neurons(n_networks, seed)
    plain_ticks = plain_ticks or ticks
    t0 = time.time()
    sims, stims = [], []
    n_seg = n_syn = 0
    for k, nn in enumerate(neurons):
        net = bion.config4_network(k, int(nn), seed)
        st = bion.soma_current(net, 200.0)
        n_seg += len(net.segments)
        n_syn += len(net.synapses)
        s = backend.Simulation(device=0, coherence=backend.COHERENCE_FAST)
        if s.load_network(net) != 0:
            raise SystemExit("plain load failed")
        del net
        s.set_stimuli(0, current=st)
        s.set_graphs(False)
        sims.append(s)
        stims.append(st)
        if (k + 1) % 64 == 0:
            print(f"# plain: {k + 1} networks loaded ({time.time() - t0:.0f}s, peak RSS {_peak_rss_gb():.1f} GB)",
                  file=sys.stderr, flush=True)
    print(f"# batch: {n_networks} networks, {n_seg} segments, {n_syn} synapses built and loaded one by one "
          f"in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    for s in sims:
        for _ in range(2):
            s.run_tick(dt)
    t0 = time.perf_counter()
    for _ in range(plain_ticks):
        for s in sims:
            s.run_tick(dt)
    plain = n_seg * plain_ticks / (time.perf_counter() - t0)
    for s in sims:
        s.close()
    del sims
    t0 = time.time()
    nets = []
    for k, nn in enumerate(neurons):
        nets.append(bion.config4_network(k, int(nn), seed))
        if (k + 1) % 64 == 0:
            print(f"# batch: {k + 1} networks built ({time.time() - t0:.0f}s, peak RSS {_peak_rss_gb():.1f} GB)",
                  file=sys.stderr, flush=True)
    b = backend.Simulation(device=0, coherence=backend.COHERENCE_FAST)
    if b.load_batch(nets) != 0:
        raise SystemExit("batch load failed")
    del nets
    print(f"# batch built and loaded in {time.time() - t0:.1f}s (peak RSS {_peak_rss_gb():.1f} GB)",
          file=sys.stderr, flush=True)
    for k, st in enumerate(stims):
        b.set_stimuli(k, current=st)
    b.set_graphs(graphs)
    tuned = b.autotune(dt, 2)
    b.run_ticks(dt, 16)
    t0 = time.perf_counter()
    b.run_ticks(dt, ticks)
    batched = n_seg * ticks / (time.perf_counter() - t0)
    sched = b.schedule().as_dict()
    b.close()
    return {"networks": n_networks, "segments": n_seg, "synapses": n_syn, "ticks": ticks,
            "plain_ticks": plain_ticks, "plain_segment_ticks_per_s": plain, "batched_segment_ticks_per_s": batched,
            "batched_vs_plain": batched / plain, "autotune_ticks": tuned, "schedule": sched}


def electrodiffusion_stress(n_cables, ticks, dt=0.025, tol=1e-6, max_sweeps=200, graphs=True):
    """The electrodiffusion stress network: n_cables unbranched 100-segment passive cables with
    linear +-20% ramps of all ions, ATP and temperature; residual-terminated Jacobi (tol mV, at
    most max_sweeps sweeps) in both potential solves; ticks replayed from captured hipGraphs.
    Its per-field error against the FP64 restatement after 500 ticks is a test
    (tests/test_residual_jacobi.py). Model bytes: 273 B per segment (dynamic state read and
    written, statics read) + 12 B per connection."""
    net = bion.config3(n_cables=n_cables, seg_per_cable=100)
    n_seg, n_conn = len(net.segments), len(net.connections)
    s = backend.Simulation(device=0, coherence=backend.COHERENCE_FAST)
    if s.load_network(net) != 0:
        raise SystemExit("electrodiffusion stress load failed")
    del net
    s.set_jacobi(tol, max_sweeps)
    s.set_graphs(graphs)
    s.run_ticks(dt, 16)
    t0 = time.perf_counter()
    s.run_ticks(dt, ticks)
    el = time.perf_counter() - t0
    s.close()
    model = 273 * n_seg + 12 * n_conn
    return {"segments": n_seg, "connections": n_conn, "ticks": ticks, "jacobi_tol_mV": tol,
            "max_sweeps": max_sweeps, "ms_per_tick": el / ticks * 1e3,
            "segment_ticks_per_s": n_seg * ticks / el, "model_bytes_per_tick": model,
            "model_GBps": model * ticks / el / 1e9}


def adjoint_cost(n_cells, ticks, every, dt=0.025, seed=0xB10C0005, graphs=True):
    """The adjoint fit: n_cells cells of 200 segments (soma + branched dendrite; Nav1.2, Kv-DR,
    K2P, NALCN, HCN1) in one batch, cable mode, a 50 pA soma current step, loss = MSE of the
    somatic V_m against the trace of the same cells with every g_max scaled by U(0.8, 1.25).
    Forward: `ticks` ticks replayed from captured hipGraphs. Adjoint: bchip_adjoint with
    checkpoints every `every` ticks (forward pass + loss, then per span a recorded recompute and
    the reverse kernel), gradients for every density and coupling scale."""
    rng = np.random.default_rng(seed)
    cells = [bion.config5_cell(seed + k, n_seg=200) for k in range(n_cells)]
    pert = [c.copy() for c in cells]
    for c in pert:
        c.segments["modifiers"]["density"][:, :5] *= rng.uniform(0.8, 1.25, (len(c.segments), 5))
    soma, off = [], 0
    for c in cells:
        soma.append(off + int(np.nonzero(c.segments["segment_type"] == 1)[0][0]))
        off += len(c.segments)
    soma = np.array(soma, np.uint32)

    def sim_of(nets):
        s = backend.Simulation(device=0, coherence=backend.COHERENCE_FAST)
        if s.load_batch([n.copy() for n in nets]) != 0:
            raise SystemExit("adjoint workload load failed")
        s.set_cable_mode(True)
        s.set_graphs(graphs)
        for k, n in enumerate(nets):
            s.set_stimuli(k, current=bion.soma_current(n, 50.0))
        return s
    s = sim_of(pert)
    _, _, _, target = s.adjoint(dt, ticks, every, soma, np.zeros((ticks, n_cells), np.float32), grads=False,
                                trace=True)
    s.close()
    s = sim_of(cells)
    s.run_ticks(dt, 32)   # capture and warm the graphs, then time from the same start state
    s.close()
    s = sim_of(cells)
    t0 = time.perf_counter()
    s.run_ticks(dt, ticks)
    fwd = time.perf_counter() - t0
    s.close()
    s = sim_of(cells)
    t0 = time.perf_counter()
    L, gd, gs = s.adjoint(dt, ticks, every, soma, target)
    adj = time.perf_counter() - t0
    s.close()
    return {"cells": n_cells, "segments": off, "ticks": ticks, "checkpoint_every": every,
            "forward_s": fwd, "adjoint_s": adj, "adjoint_vs_forward": adj / fwd, "loss": L,
            "grad_density_norm": float(np.linalg.norm(gd)), "grad_scale_norm": float(np.linalg.norm(gs))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--neurons", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compat-ticks", type=int, default=3,
                    help="ticks timed in COMPAT coherence after the FAST ticks (0: skip)")
    ap.add_argument("--batch-networks", type=int, default=512,
                    help="networks in the batched-vs-plain comparison (512: the batched sweep's ~20M segments)")
    ap.add_argument("--batch-ticks", type=int, default=1000, help="ticks of the batched run")
    ap.add_argument("--plain-ticks", type=int, default=50, help="ticks of the one-network-at-a-time run")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-vs-plain comparison")
    ap.add_argument("--stress-cables", type=int, default=10_000,
                    help="100-segment cables in the electrodiffusion stress leg (0: skip)")
    ap.add_argument("--stress-ticks", type=int, default=500)
    ap.add_argument("--stress-only", action="store_true", help="only the electrodiffusion stress leg")
    ap.add_argument("--adjoint-cells", type=int, default=64)
    ap.add_argument("--adjoint-ticks", type=int, default=20000)
    ap.add_argument("--adjoint-every", type=int, default=500)
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint/forward cost")
    ap.add_argument("--autotune-ticks", type=int, default=0,
                    help="live ticks per schedule candidate of the autotuner before the tick leg (0: defaults)")
    ap.add_argument("--schedule", default="",
                    help="launch schedule as field=value[,field=value] (replaces the autotuner's search)")
    ap.add_argument("--step1-chunks", type=int, default=0,
                    help="override the schedule's step1_chunks (1: steps 1 and 2.5 on one stream)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="tick by tick instead of captured hipGraphs in the batched, stress and adjoint "
                         "legs (for kernel tracing)")
    ap.add_argument("--cpu-full-ticks", type=int, default=0,
                    help="only time the CPU tick on the full workload: 2 warm-up + this many ticks")
    ap.add_argument("--cpu-calibration", action="store_true",
                    help="only time the oracle against the compiled reference on the same ticks")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 1 GiB device-to-device copies (the reference kernel for "
                         "scaling FETCH_SIZE/WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "copies": 3}), flush=True)
        return
    if args.cpu_full_ticks > 0:
        print(json.dumps(cpu_full(args.neurons, args.cpu_full_ticks)), flush=True)
        return
    if args.cpu_calibration:
        print(json.dumps(cpu_calibration()), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; runs on more than one are not launched",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.batch_only:
        print(json.dumps(batched_vs_plain(args.batch_networks, args.batch_ticks, graphs=not args.no_graphs,
                                          plain_ticks=args.plain_ticks)), flush=True)
        return
    if args.stress_only:
        print(json.dumps(electrodiffusion_stress(args.stress_cables, args.stress_ticks, graphs=not args.no_graphs)),
              flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_cost(args.adjoint_cells, args.adjoint_ticks, args.adjoint_every,
                                      graphs=not args.no_graphs)), flush=True)
        return
    t0 = time.time()
    net, stims = build_workload(args.neurons, args.warmup + args.steps)
    print(f"# built {len(net.segments)} segments, {len(net.synapses)} synapses in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    sim = backend.Simulation(device=0, coherence=backend.COHERENCE_FAST)
    t0 = time.time()
    rc = sim.load_network(net)
    if rc != 0:
        raise SystemExit(f"load failed: {rc}")
    print(f"# loaded in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    tuned = 0
    if args.schedule:
        sim.set_schedule(**{k: int(v) for k, v in (f.split("=") for f in args.schedule.split(","))})
    elif args.autotune_ticks > 0:   # the schedule search runs live ticks of this network
        tuned = sim.autotune(0.025, args.autotune_ticks)
        print(f"# autotuned in {tuned} ticks: {sim.schedule().as_dict()}", file=sys.stderr, flush=True)
    if args.step1_chunks > 0:
        sim.set_schedule(step1_chunks=args.step1_chunks)
    nt_in_live = bool(np.any(net.segments["intra_nt"].view(np.uint32)))
    algo = algorithmic_bytes(net, len(net.synapses) == 0, nt_in_live)
    tick_bytes = tick_bytes_model(net)
    n_seg = len(net.segments)
    del net
    for t in range(args.warmup):
        sim.upload_stimuli(stims[t])
        sim.run_tick(0.025)
    sim.set_phase_timing(True)
    phase = np.zeros(len(PHASES))
    t0 = time.perf_counter()
    for t in range(args.warmup, args.warmup + args.steps):
        sim.upload_stimuli(stims[t])
        sim.run_tick(0.025)
        phase += np.array(sim.stats().last_phase_ms[:])
    el = time.perf_counter() - t0
    sim.set_phase_timing(False)
    compat = None
    if args.compat_ticks > 0:
        # COMPAT coherence, the drop-in default: the host records are uploaded before and
        # refreshed after every tick (the PCIe-inclusive rate of a caller that hands over host
        # buffers each tick)
        sim.set_coherence(backend.COHERENCE_COMPAT)
        sim.run_tick(0.025)
        t0 = time.perf_counter()
        for _ in range(args.compat_ticks):
            sim.run_tick(0.025)
        c_el = time.perf_counter() - t0
        compat = {"ticks": args.compat_ticks, "ms_per_tick": c_el / args.compat_ticks * 1e3,
                  "segment_ticks_per_s": n_seg * args.compat_ticks / c_el}
    sched = sim.schedule().as_dict()
    sim.close()   # its host records and device state are not needed by the other legs
    del sim, stims
    ms_step = el / args.steps * 1e3
    value = n_seg * args.steps / el
    mean = phase / args.steps
    dom = int(np.argmax(mean))
    name = PHASES[dom]
    achieved = algo[name] / (mean[dom] * 1e-3) / 1e9
    traffic = measured_traffic(name, args.neurons)
    valu = measured_valu(name, args.neurons)
    valu_frac = (valu * VALU_CYCLES / VALU_SIMDS / CLOCK_HZ) / (mean[dom] * 1e-3) if valu else None
    tick_gbs = tick_bytes / (ms_step * 1e-3) / 1e9
    out = {
        "metric": "segment-ticks/s @3.2M seg (HBM roofline %); batched vs plain; adjoint/fwd cost",
        "value": value, "unit": "segment-ticks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32 (reference ABI mixed precision)", "data": "synthetic",
        "config": {"workload": "plain tick, 100k neurons x 32 segments tree = 3.2M segments, 12 modifiers/segment, "
                               "10M AMPA+NMDA/GABA-A synapses, Poisson 5 Hz somatic pulses, dt 0.025 ms",
                   "neurons": args.neurons, "segments": n_seg},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": algo[name], "launch_ms": float(mean[dom]),
                     # what limits it: the share of the launch its VALU instructions occupy every
                     # SIMD (SQ_INSTS_VALU x 4 cycles / 1024 SIMDs / clock / duration)
                     "limiter": "valu" if valu_frac is not None and valu_frac > achieved / HBM_PEAK_GBS else "hbm",
                     "valu_frac": valu_frac, "valu_insts": valu},
        "tick_roofline": {"model_bytes_per_tick": tick_bytes, "achieved": tick_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": tick_gbs / HBM_PEAK_GBS},
        "phase_ms": {PHASES[k]: float(mean[k]) for k in range(len(PHASES))},
        "schedule": sched, "autotune_ticks": tuned,
    }
    if compat:
        out["compat"] = compat
    if args.batch_networks > 0:
        out["batched"] = batched_vs_plain(args.batch_networks, args.batch_ticks, plain_ticks=args.plain_ticks)
    if args.stress_cables > 0:
        out["electrodiffusion"] = electrodiffusion_stress(args.stress_cables, args.stress_ticks)
    if args.adjoint_cells > 0:
        out["adjoint"] = adjoint_cost(args.adjoint_cells, args.adjoint_ticks, args.adjoint_every)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
