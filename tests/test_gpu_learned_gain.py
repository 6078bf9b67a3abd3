"""What the shipped learned preconditioner gains over the plain method.

Two p' systems per size: the d = 1 cavity operator with a seeded mean-free
right-hand side, and the operator and right-hand side of a 320^2 cavity at
Re = 2000 after three SIMPLE iterations (non-uniform response coefficients;
a case neither trained on nor used for model selection). Every method starts
from x0 = 0 and must reach the same FP64 true residual, |b - A x| <= 1e-8 |b|.

Flexible CG with the shipped network (cfdapp_amd/weights: the V-cycle
parametrisation with Neumann padding and the coarse tail, as selected by the
training run on held-out systems) must need at most a twentieth of the
iterations of Jacobi-PCG (the method without it) on the d = 1 systems and at
most a fifth on the SIMPLE operator, and fewer iterations than the
zero-padded additive multilevel network (the other parametrisation the
trainer offers). Iteration counts and device times are printed side by side.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_pair(d, weights, options):
    import cfdapp_amd as cfd
    jac = cfd.LinearSolverSettings(0.0, 1e-8, 400000, "CG", "Jacobi", check_interval=64)
    fcg = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
    d.set_learned_weights(weights, options=options)
    rj = d.solve_pressure_rhs(jac, None, download=True)
    rf = d.solve_pressure_rhs(fcg, None, download=True)
    return rj, rf


def test_shipped_weights_are_the_v_cycle_network():
    import json
    import os

    import cfdapp_amd as cfd
    from cfdapp_amd.learned import VCycle
    opts = cfd.load_learned_options()
    assert opts["padding"] == "neumann" and opts["tail"]["levels"] > 0
    w = cfd.load_learned_weights()
    assert np.isfinite(w).all()
    with open(os.path.splitext(cfd.WEIGHTS_PATH)[0] + ".json") as f:
        rec = json.load(f)
    assert rec["parametrisation"] == "VCycle"
    np.testing.assert_array_equal(w, VCycle.weights(rec["theta"]))


@pytest.mark.parametrize("n", [256, 512])
def test_learned_gain_standalone(n):
    import cfdapp_amd as cfd
    from cfdapp_amd import learned
    w, opts = cfd.load_learned_weights(), cfd.load_learned_options()
    tr = learned.evaluate(w, n, options=opts)
    ml = learned.evaluate(learned.structured_init(0, noise=0.0), n)
    print(f"{n}^2 d=1: shipped {tr['fcg_iterations']} it {tr['fcg_ms']:.1f} ms, "
          f"additive multilevel {ml['fcg_iterations']} it {ml['fcg_ms']:.1f} ms, "
          f"Jacobi-PCG {tr['jacobi_iterations']} it {tr['jacobi_ms']:.1f} ms", flush=True)
    for r in (tr, ml):
        assert r["fcg_status"] == "Converged" and not r["fcg_fallback"]
        assert r["jacobi_status"] == "Converged"
        assert r["fcg_true_rel"] <= 1e-8
        # Jacobi-PCG stops on its recurrence residual (reference semantics)
        assert r["jacobi_true_rel"] <= 1.05e-8
    assert 20 * tr["fcg_iterations"] <= tr["jacobi_iterations"]
    assert tr["fcg_iterations"] < ml["fcg_iterations"]


def test_learned_gain_simple_operator():
    import cfdapp_amd as cfd
    from cfdapp_amd import learned
    d = learned.make_case("cavity", 320, 2000.0, 1.0)
    try:
        learned.assemble_pprime(d, 3)
        b = d.download("PC_RHS")
        rj, rf = _solve_pair(d, cfd.load_learned_weights(), cfd.load_learned_options())
        _, rm = _solve_pair(d, learned.structured_init(0, noise=0.0), {"padding": "zero"})
        print(f"320^2 Re 2000 SIMPLE operator: shipped {rf.iterations} it {rf.device_ms:.1f} ms, "
              f"additive multilevel {rm.iterations} it {rm.device_ms:.1f} ms, "
              f"Jacobi-PCG {rj.iterations} it {rj.device_ms:.1f} ms", flush=True)
        bn = np.linalg.norm(b)
        for r in (rj, rf, rm):
            assert r.status == 0
        for r in (rf, rm):
            assert r.true_residual <= 1e-8 * bn * (1 + 1e-9)
        assert rj.true_residual <= 1.05e-8 * bn
        assert not rf.fallback_used
        # the two solutions agree to what the tolerance allows
        assert np.linalg.norm(rf.solution - rj.solution) <= 1e-4 * np.linalg.norm(rj.solution)
        assert 5 * rf.iterations <= rj.iterations
        assert rf.iterations < rm.iterations
    finally:
        d.close()
