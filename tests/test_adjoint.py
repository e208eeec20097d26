"""The adjoint of the cable-mode tick (bchip_adjoint) on the GPU.

The workload: two branched cells (200 and 120 segments; Nav1.2, Kv-DR, K2P, NALCN, HCN1 in every
segment) loaded as one batch, a current step into each soma, per-connection coupling scales
drawn in [0.5, 1.5], and the loss mean over 48 ticks and both cells of (V_soma - target)^2.

- the cable-mode forward is bit-identical to the oracle's cable mode;
- gradients with respect to densities and coupling scales equal central differences (h = 1e-3
  relative) of the oracle's FP64 cable restatement: per parameter for the somatic channels and
  sampled dendritic ones, and along whole-type, all-scale and random directions;
- directional central differences of the GPU's own forward agree to the float forward's noise;
- the checkpoint interval changes neither the loss nor the gradients (bit for bit), and the run
  ends in the state bchip_run_ticks reaches;
- a torch.autograd.Function over the C ABI (test-side only) gives the same gradients, and a step
  against them lowers the loss.
"""
import numpy as np
import pytest
import torch

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025
TICKS = 48
SIZES = (200, 120)
H_REL = 1e-3


def _cells():
    return [bion.config5_cell(51 + k, n_seg=n) for k, n in enumerate(SIZES)]


def _scales(cells, seed=7):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.5, 1.5, len(c.connections)).astype(np.float32) for c in cells]


def _stim(c, k):
    return H.soma_clamp(c, 40.0 + 30.0 * k)


def _soma(cells):
    out, off = [], 0
    for c in cells:
        out.append(off + int(np.nonzero(c.segments["segment_type"] == 1)[0][0]))
        off += len(c.segments)
    return np.array(out, np.uint32)


def _target():
    t = np.arange(TICKS, dtype=np.float32)
    return np.stack([-58.0 + 0.2 * t, -50.0 - 0.1 * t], axis=1).astype(np.float32)


def _gpu(cells, scales):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([c.copy() for c in cells]) == 0
    sim.set_cable_mode(True)
    sim.set_coupling_scale(np.concatenate(scales))
    for k, c in enumerate(cells):
        sim.set_stimuli(k, current=_stim(c, k))
    return sim


def _fp64_loss(cells, scales):
    """The loss of the batch from the FP64 restatement (one cell at a time)."""
    tgt = _target()
    tot = 0.0
    for k, c in enumerate(cells):
        soma = int(np.nonzero(c.segments["segment_type"] == 1)[0][0])
        tot += H.Cable64(c.segments, c.connections, scale=scales[k], current=_stim(c, k)).loss(
            DT, TICKS, [soma], tgt[:, k:k + 1])
    return tot / len(cells)


def _perturbed(cells, scales, dens=None, scale=None):
    """Copies with density entries {(cell, seg, slot): factor} and scale entries {(cell, conn): factor}."""
    cs = [c.copy() for c in cells]
    ss = [s.copy() for s in scales]
    for (k, i, j), f in (dens or {}).items():
        cs[k].segments["modifiers"]["density"][i, j] *= f
    for (k, e), f in (scale or {}).items():
        ss[k][e] *= f
    return cs, ss


def _adjoint(cells, scales, K=16, grads=True):
    sim = _gpu(cells, scales)
    out = sim.adjoint(DT, TICKS, K, _soma(cells), _target(), grads=grads)
    sim.close()
    return out


def _split(cells, gd, gs):
    """Per-cell views of the host-order gradients."""
    out, so, co = [], 0, 0
    for c in cells:
        n, m = len(c.segments), len(c.connections)
        out.append((gd[so:so + n], gs[co:co + m]))
        so += n
        co += m
    return out


def test_cable_forward_equals_oracle_cable_mode():
    cells, scales = _cells(), _scales(_cells())
    sim = _gpu(cells, scales)
    sim.run_ticks(DT, TICKS)
    for k, c in enumerate(cells):
        seg, _, _ = sim.batch_read(k)
        o = H.OracleSim(c.copy())
        o.set_cable(True, scales[k])
        for _ in range(TICKS):
            o.step(DT, H.Stimuli(current=_stim(c, k)))
        assert H.first_mismatch(o.net.segments, seg) is None, (k, H.first_mismatch(o.net.segments, seg))
    sim.close()


def test_gradients_equal_fp64_central_differences():
    cells, scales = _cells(), _scales(_cells())
    L, gd, gs = _adjoint(cells, scales)
    L64 = _fp64_loss(cells, scales)
    print(f"loss GPU {L:.9g} FP64 {L64:.9g}", flush=True)
    assert abs(L - L64) < 1e-4 * L64
    per = _split(cells, gd, gs)
    rng = np.random.default_rng(3)

    def fd(dens=None, scale=None):
        lp = _fp64_loss(*_perturbed(cells, scales, {q: 1 + H_REL for q in dens or {}}, {q: 1 + H_REL for q in scale or {}}))
        lm = _fp64_loss(*_perturbed(cells, scales, {q: 1 - H_REL for q in dens or {}}, {q: 1 - H_REL for q in scale or {}}))
        return (lp - lm) / (2 * H_REL)   # d loss / d log(parameter) along the direction

    checks = []
    for k, c in enumerate(cells):
        soma = int(np.nonzero(c.segments["segment_type"] == 1)[0][0])
        dens = c.segments["modifiers"]["density"]
        gdk, gsk = per[k]
        params = [(soma, j) for j in range(5)]
        params += [(int(i), int(j)) for i, j in zip(rng.integers(1, len(c.segments), 6), rng.integers(0, 5, 6))]
        for i, j in params:
            checks.append((f"cell {k} seg {i} slot {j}", gdk[i, j] * dens[i, j], fd(dens=[(k, i, j)])))
        for e in [0, 1] + list(rng.integers(2, len(c.connections), 5)):
            e = int(e)
            checks.append((f"cell {k} conn {e} scale", gsk[e] * scales[k][e], fd(scale=[(k, e)])))
        for j in range(5):   # the whole type j over the cell
            q = [(k, i, j) for i in range(len(c.segments))]
            checks.append((f"cell {k} all slot {j}", float((gdk[:, j] * dens[:, j]).sum()), fd(dens=q)))
        checks.append((f"cell {k} all scales", float((gsk * scales[k]).sum()),
                       fd(scale=[(k, e) for e in range(len(c.connections))])))
        # a random direction over every density and scale of the cell: log p += h * r, r = +-1
        rd = rng.choice([-1.0, 1.0], size=(len(c.segments), 5))
        rs = rng.choice([-1.0, 1.0], size=len(c.connections))
        a = float((gdk[:, :5] * dens[:, :5] * rd).sum() + (gsk * scales[k] * rs).sum())
        fp = _fp64_loss(*_perturbed(cells, scales, {(k, i, j): 1 + H_REL * rd[i, j] for i in range(len(c.segments)) for j in range(5)},
                                    {(k, e): 1 + H_REL * rs[e] for e in range(len(c.connections))}))
        fm = _fp64_loss(*_perturbed(cells, scales, {(k, i, j): 1 - H_REL * rd[i, j] for i in range(len(c.segments)) for j in range(5)},
                                    {(k, e): 1 - H_REL * rs[e] for e in range(len(c.connections))}))
        checks.append((f"cell {k} random direction", a, (fp - fm) / (2 * H_REL)))
    big = max(abs(f) for _, _, f in checks)
    bad = []
    for name, a, f in checks:
        err = abs(a - f)
        print(f"{name:28s} adjoint {a: .6e} fd {f: .6e} err {err:.2e}", flush=True)
        if err > 1e-3 * abs(f) + 1e-6 * big:
            bad.append(name)
    assert not bad, bad


def test_gradients_match_gpu_forward_differences():
    cells, scales = _cells(), _scales(_cells())
    L, gd, gs = _adjoint(cells, scales)
    per = _split(cells, gd, gs)
    for j in (0, 3, 4):   # Nav1.2, NALCN and HCN1 of both cells together
        dens = {(k, i, j): None for k, c in enumerate(cells) for i in range(len(c.segments))}
        lp = _adjoint(*_perturbed(cells, scales, {q: 1 + H_REL for q in dens}), grads=False)[0]
        lm = _adjoint(*_perturbed(cells, scales, {q: 1 - H_REL for q in dens}), grads=False)[0]
        f = (lp - lm) / (2 * H_REL)
        a = sum(float((per[k][0][:, j] * c.segments["modifiers"]["density"][:, j]).sum()) for k, c in enumerate(cells))
        print(f"slot {j}: adjoint {a:.6e} GPU fd {f:.6e}", flush=True)
        assert abs(a - f) < 2e-2 * abs(f)


def test_checkpoint_interval_recording_and_end_state():
    """Checkpoint intervals, and recorded iterates vs substeps run again in the reverse pass."""
    cells, scales = _cells(), _scales(_cells())
    ref = None
    for K, rec in ((TICKS, True), (16, True), (7, False), (1, True), (16, False)):
        sim = _gpu(cells, scales)
        if not rec:
            sim.set_adjoint_memory(0)
        L, gd, gs = sim.adjoint(DT, TICKS, K, _soma(cells), _target())
        ends = [sim.batch_read(k) for k in range(len(cells))]
        assert sim.tick_count() == TICKS
        sim.close()
        if ref is None:
            ref = (L, gd, gs)
            plain = _gpu(cells, scales)
            plain.run_ticks(DT, TICKS)
            for k in range(len(cells)):
                seg, _, spk = plain.batch_read(k)
                assert H.first_mismatch(seg, ends[k][0]) is None
                np.testing.assert_array_equal(spk, ends[k][2])
            plain.close()
        else:
            assert L == ref[0]
            np.testing.assert_array_equal(gd, ref[1])
            np.testing.assert_array_equal(gs, ref[2])


class CableLoss(torch.autograd.Function):
    """loss(density, scale) of the batch through bchip_adjoint; density: [segments, 5]."""

    @staticmethod
    def forward(ctx, density, scale, cells):
        cs = [c.copy() for c in cells]
        so = co = 0
        ss = []
        for c in cs:
            n, m = len(c.segments), len(c.connections)
            c.segments["modifiers"]["density"][:, :5] = density[so:so + n].detach().cpu().numpy()
            ss.append(scale[co:co + m].detach().cpu().numpy().astype(np.float32))
            so += n
            co += m
        L, gd, gs = _adjoint(cs, ss)
        ctx.save_for_backward(torch.from_numpy(gd[:, :5].copy()), torch.from_numpy(gs.copy()))
        return torch.tensor(L, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gd, gs = ctx.saved_tensors
        return g * gd, g * gs, None


def test_autograd_function_and_descent():
    cells, scales = _cells(), _scales(_cells())
    dens = torch.tensor(np.concatenate([c.segments["modifiers"]["density"][:, :5] for c in cells]),
                        dtype=torch.float64, requires_grad=True)
    sc = torch.tensor(np.concatenate(scales), dtype=torch.float64, requires_grad=True)
    L = CableLoss.apply(dens, sc, cells)
    L.backward()
    _, gd, gs = _adjoint(cells, scales)
    np.testing.assert_array_equal(dens.grad.numpy(), gd[:, :5])
    np.testing.assert_array_equal(sc.grad.numpy(), gs)
    # one step along the negative gradient in log-parameters lowers the loss
    with torch.no_grad():
        gl = torch.cat([(dens.grad * dens).flatten(), sc.grad * sc])
        step = 1e-2 / float(gl.abs().max())
        d2 = dens * torch.exp(-step * dens.grad * dens)
        s2 = sc * torch.exp(-step * sc.grad * sc)
    L2 = CableLoss.apply(d2, s2, cells)
    print(f"loss {L.item():.6g} -> {L2.item():.6g}", flush=True)
    assert L2.item() < L.item()


def test_adjoint_rejects_unsupported_runs():
    cells = _cells()
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([c.copy() for c in cells]) == 0
    with pytest.raises(RuntimeError):   # cable mode off
        sim.adjoint(DT, 4, 2, _soma(cells), np.zeros((4, 2), np.float32))
    sim.close()
    net = bion.config4_network(3, 5)   # synapses and channels outside the differentiated set
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_network(net.copy()) == 0
    with pytest.raises(ValueError):
        sim.set_cable_mode(True)
    sim.close()


def test_long_horizon_gradients_with_config5_checkpoints():
    """2000 ticks (50 ms) with checkpoints every 500 ticks, the adjoint workload's interval: the
    loss equals the FP64 restatement's and the gradient's directional derivatives (each channel
    type over both cells, every coupling scale, a random direction per cell) equal FP64 central
    differences (h = 1e-3 relative) within 1e-3 relative."""
    T, K = 2000, 500
    cells, scales = _cells(), _scales(_cells())
    t = np.arange(T, dtype=np.float32)
    tgt = np.stack([-60.0 + 4.0 * np.sin(t / 300.0), -55.0 - 3.0 * np.cos(t / 250.0)], axis=1).astype(np.float32)
    sim = _gpu(cells, scales)
    L, gd, gs = sim.adjoint(DT, T, K, _soma(cells), tgt)
    sim.close()

    def loss64(cs, ss):
        tot = 0.0
        for k, c in enumerate(cs):
            soma = int(np.nonzero(c.segments["segment_type"] == 1)[0][0])
            tot += H.Cable64(c.segments, c.connections, scale=ss[k], current=_stim(c, k)).loss(
                DT, T, [soma], tgt[:, k:k + 1])
        return tot / len(cs)

    L64 = loss64(cells, scales)
    print(f"loss GPU {L:.9g} FP64 {L64:.9g}", flush=True)
    assert abs(L - L64) < 1e-3 * L64
    per = _split(cells, gd, gs)
    rng = np.random.default_rng(11)
    checks = []

    def fd(dens, scale):
        lp = loss64(*_perturbed(cells, scales, {q: 1 + H_REL * r for q, r in dens.items()},
                                {q: 1 + H_REL * r for q, r in scale.items()}))
        lm = loss64(*_perturbed(cells, scales, {q: 1 - H_REL * r for q, r in dens.items()},
                                {q: 1 - H_REL * r for q, r in scale.items()}))
        return (lp - lm) / (2 * H_REL)

    for j in range(5):
        q = {(k, i, j): 1.0 for k, c in enumerate(cells) for i in range(len(c.segments))}
        a = sum(float((per[k][0][:, j] * c.segments["modifiers"]["density"][:, j]).sum()) for k, c in enumerate(cells))
        checks.append((f"slot {j} over both cells", a, fd(q, {})))
    q = {(k, e): 1.0 for k, c in enumerate(cells) for e in range(len(c.connections))}
    a = sum(float((per[k][1] * scales[k]).sum()) for k in range(len(cells)))
    checks.append(("every coupling scale", a, fd({}, q)))
    for k, c in enumerate(cells):
        dens = c.segments["modifiers"]["density"]
        rd = rng.choice([-1.0, 1.0], size=(len(c.segments), 5))
        rs = rng.choice([-1.0, 1.0], size=len(c.connections))
        a = float((per[k][0][:, :5] * dens[:, :5] * rd).sum() + (per[k][1] * scales[k] * rs).sum())
        f = fd({(k, i, j): rd[i, j] for i in range(len(c.segments)) for j in range(5)},
               {(k, e): rs[e] for e in range(len(c.connections))})
        checks.append((f"cell {k} random direction", a, f))
    big = max(abs(f) for _, _, f in checks)
    bad = []
    for name, a, f in checks:
        err = abs(a - f)
        print(f"{name:28s} adjoint {a: .6e} fd {f: .6e} err {err:.2e}", flush=True)
        if err > 1e-3 * abs(f) + 1e-6 * big:
            bad.append(name)
    assert not bad, bad
