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
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bioneuroncore_amd import backend, bion  # noqa: E402

HBM_PEAK_GBS = 8000.0
PHASES = ("k_modifiers", "k_diffusion", "k_syn_pre", "k_syn_post", "k_syn_plast", "k_syn_rrp")


def algorithmic_bytes(net, stats):
    """Compulsory HBM bytes per launch of each kernel: every dynamic field it owns read once and
    written once, every static parameter it needs read once (P64 = the ABI's precision)."""
    n = len(net.segments)
    nc = len(net.connections)
    ns = len(net.synapses)
    types = net.segments["modifiers"]["type"]
    count = net.segments["modifier_count"]
    # dynamic gate states per modifier type (written) and parameter states read
    gate_w = {142: 2, 141: 2, 156: 2, 150: 3, 112: 1, 140: 2, 129: 1, 103: 1, 104: 1, 107: 0, 7: 0, 133: 0, 123: 1}
    param_r = {142: 1, 141: 1, 156: 2, 150: 1, 112: 5, 140: 0, 129: 5, 103: 3, 104: 2, 107: 2, 7: 0, 133: 1, 123: 3}
    gates = params = 0
    for t in np.unique(types[types > 0]):
        k = int((types == t).sum())
        gates += k * gate_w.get(int(t), 0)
        params += k * param_r.get(int(t), 0)
    nmod = int(count.sum())
    seg_dyn = 8 + 7 * 8 + 7 * 8 + 4            # V, intra, extra, ATP
    mod = {
        "k_modifiers": n * (2 * seg_dyn + 8 + 4 + 8) + nmod * (1 + 8) + gates * 16 + params * 8,
        "k_diffusion": n * (2 * (8 + 7 * 8 + 7 * 8) + 8 + 4 + 4 + 4 + 8 + 2 * 4 * 3) + nc * 12,
        "k_syn_pre": ns * (4 + 4 + 4 + 4 + 4 + 2 * 4 + 1 + 2 * 1 + 4) + 64 * 0,
        # ordered pass: segment V, ions in/out, ibuf, cap, transmitter out; per synapse flags, mask,
        # one cleft transmitter, receptor count, post area, receptor current out
        "k_syn_post": n * (2 * (8 + 7 * 8) + 7 * 8 + 8 + 4 + 15 * 4) + ns * (1 + 2 + 4 + 1 + 4 + 4),
        # flags, current, mask, tau, pp1 read; spine Ca, one cleft transmitter, BCM, CaMKII,
        # calcineurin, weight read+written; depression written
        "k_syn_plast": ns * (1 + 4 + 2 + 4 + 4 + 6 * 2 * 4 + 4),
        "k_syn_rrp": ns * (4 + 4 + 4 + 2 * 4 + 4),
    }
    rec = net.synapses["post_receptor_count"].astype(np.int64).sum()
    mod["k_syn_post"] += int(rec) * (1 + 8 + 2 * 2 * 8)
    return mod


def build_workload(neurons, ticks, seed=0xB10C0002):
    net = bion.config2(n_neurons=neurons, seed=seed)
    pulses = bion.poisson_soma_pulses(net, 5.0, ticks, 0.025, 1000.0, seed + 1)
    stims = [backend.make_stimuli(current=p[:256]) for p in pulses]
    return net, stims


def cpu_baseline(budget_s=12.0, neurons=250):
    """Reference CPU tick on a bounded sample of the same workload (250 neurons = 8000 segments,
    25k synapses), one thread, ticks until the time budget is used."""
    net, stims = build_workload(neurons, 4000, seed=0xB10C0102)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    P = C.c_void_p
    kind = "reference" if H.ref_available() else "port"
    if kind == "reference":
        lib = C.CDLL(H.REF_SO)
        lib.core_run_tick.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, C.c_uint8, P, C.c_float] + [P] * 12
        tick = np.zeros(1, np.uint64)
        mech = np.zeros(4, np.uint8)
        mc = np.zeros(1, np.uint32)

        def step(t):
            st = H.Stimuli(current=stims[t]["current_stim"][0][:stims[t]["current_stim_count"][0]])
            lib.core_run_tick(net.segments.ctypes.data, len(net.segments), net.connections.ctypes.data,
                              len(net.connections), net.synapses.ctypes.data, len(net.synapses), 0,
                              tick.ctypes.data, 0.025, *st.args(), mech.ctypes.data, mc.ctypes.data)
    else:
        sim = H.OracleSim(net)

        def step(t):
            sim.step(0.025, H.Stimuli(current=stims[t]["current_stim"][0][:stims[t]["current_stim_count"][0]]))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--neurons", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
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
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one GPU; multi-GPU sharding is described in DESIGN.md, "
                                   "not launched", "n_gpus": args.gpus}))
        sys.exit(2)
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
    algo = algorithmic_bytes(net, sim.stats())
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
    ms_step = el / args.steps * 1e3
    value = n_seg * args.steps / el
    mean = phase / args.steps
    dom = int(np.argmax(mean))
    name = PHASES[dom]
    achieved = algo[name] / (mean[dom] * 1e-3) / 1e9
    out = {
        "metric": "segment-ticks/s @3.2M seg (HBM roofline %); batched vs plain; adjoint/fwd cost",
        "value": value, "unit": "segment-ticks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32 (reference ABI mixed precision)", "data": "synthetic",
        "config": {"workload": "plain tick, 100k neurons x 32 segments tree = 3.2M segments, 12 modifiers/segment, "
                               "10M AMPA+NMDA/GABA-A synapses, Poisson 5 Hz somatic pulses, dt 0.025 ms",
                   "neurons": args.neurons, "segments": n_seg},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "phase_ms": {PHASES[k]: float(mean[k]) for k in range(len(PHASES))},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
