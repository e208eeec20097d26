"""The oracle's residual-terminated Jacobi (an extension) and its FP64 restatement of passive
networks, pinned by: the residual rule with tol -> 0 and 50 sweeps reproducing the reference's
fixed sweeps bit for bit; per-cable conservation of sum(vol * C) for every intracellular ion
in FP64 (the intracellular Nernst-Planck step moves flux/vol_a out of a and flux/vol_b into b,
include/diffusion_kernel.h:202-203); and agreement of the FP64 and mixed-precision ticks
within the mixed precision's rounding."""
import numpy as np

from bioneuroncore_amd import bion
import fp64
import harness as H

DT = 0.025


def test_residual_rule_with_zero_tolerance_is_the_reference():
    for net in (bion.config1(seed=5), bion.config3(n_cables=3, seg_per_cable=40, seed=6)):
        a = H.OracleSim(net.copy())
        b = H.OracleSim(net.copy())
        fp64.lib().oracle_set_jacobi(b.ctx, 1e-300, 50)
        stim = H.Stimuli(current=H.soma_clamp(net, 200.0))
        for _ in range(25):
            a.step(DT, stim)
            b.step(DT, stim)
        assert H.first_mismatch(a.net.segments, b.net.segments) is None


def test_fp64_conserves_ions_per_cable():
    net = bion.config3(n_cables=4, seg_per_cable=60, seed=8)
    s = net.segments
    r = s["radius"].astype(np.float64)
    l = s["length"].astype(np.float64)
    sphere = np.isin(s["segment_type"], (1, 3, 4))   # soma-like segments are spheres
    vol = np.where(sphere, 4.0 / 3.0 * np.pi * r ** 3, np.pi * r * r * l)
    cable = np.repeat(np.arange(4), 60)
    before = np.stack([np.bincount(cable, vol * s["intra_ions"][:, k]) for k in range(7)])
    out = fp64.fp64_run(net, 200, DT, 1e-6, 200)
    after = np.stack([np.bincount(cable, vol * out["intra"][:, k]) for k in range(7)])
    np.testing.assert_allclose(after, before, rtol=1e-12, atol=1e-18)


def test_fp64_agrees_with_mixed_precision():
    net = bion.config3(n_cables=3, seg_per_cable=100, seed=9)
    m = H.OracleSim(net.copy())
    fp64.lib().oracle_set_jacobi(m.ctx, 1e-6, 200)
    for _ in range(200):
        m.step(DT)
    ref = fp64.fp64_run(net, 200, DT, 1e-6, 200)
    err = fp64.field_errors(m.net.segments, ref)
    print(err)
    assert err["V_m"] < 1e-4 and err["ATP"] < 1e-4 and err["temperature"] < 1e-6
    assert max(err[f"intra[{k}]"] for k in range(7)) < 1e-6
    assert max(err[f"extra[{k}]"] for k in range(7)) < 1e-9


def test_fp64_build_agrees_with_passive_restatement():
    """The oracle's FP64 build (every float quantity in double) against the independent FP64
    restatement of passive networks on 100-segment cables with +-20% ramps, 100 ticks: V_m and
    every ion to 1e-8 relative; ATP to float storage (the build keeps the record's float ATP)."""
    import f64full
    net = bion.config3(n_cables=4, seg_per_cable=100)
    ref = fp64.fp64_run(net.copy(), 100, 0.025)
    b = net.copy()
    f64full.run(b, 100, 0.025)
    err = fp64.field_errors(b.segments, ref)
    print({k: f"{v:.1e}" for k, v in err.items()})
    assert err["V_m"] < 1e-8 and err["temperature"] < 1e-8
    assert max(v for k, v in err.items() if k.startswith(("intra", "extra"))) < 1e-8
    assert err["ATP"] < 1e-5
