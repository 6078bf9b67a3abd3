"""GPU parity of the learned preconditioner's network and its flexible-CG solve.

The MFMA forward and backward kernels are compared with the CPU oracle
(oracle/unet.py) run with the device's rounding points (forward_device /
backward_device: BF16 weights, FP32 activations split into hi + lo BF16 at the
MFMA, BF16 operands in the weight gradients). What remains is the FP32
accumulation order: the bounds are 1e-4 relative (L2) on every activation and
2e-3 on every layer's weight gradient (whose BF16 operands can round the other
way). A layout or indexing mistake gives O(1) errors. The exact FP64 network
with the same BF16 weights is checked at 1e-3 as a bound on the emulation.
"""
import numpy as np
import pytest

from oracle import unet

torch = pytest.importorskip("torch")
import cfdapp_amd as ca  # noqa: E402

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    torch.cuda.synchronize()
    return t


def _w64(w):
    return {k: v.astype(np.float64) for k, v in w.items()}


@pytest.mark.parametrize("n,h,w", [(2, 32, 64), (1, 24, 40), (1, 64, 96)])
def test_forward_matches_oracle(n, h, w):
    print(f"forward n={n} h={h} w={w}", flush=True)
    wts = unet.init_weights(1)
    net = ca.UnetNet(unet.flatten(wts))
    x = np.random.default_rng(h * w).standard_normal((n, h, w)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty_like(xd)
    net.forward(xd, yd, n, h, w)
    y = yd.cpu().numpy()
    yb, tape = unet.forward_device(_w64(wts), x[..., None])
    errs = {name: _rel(net.tensor(name), tape[name].ravel()) for name in ca.UnetNet.TENSORS[:17]}
    print("activation errors", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert _rel(y, yb[..., 0]) < 1e-4
    for name, err in errs.items():
        assert err < 1e-4, name
    wb = {k: unet.bf(v) for k, v in _w64(wts).items()}
    y64 = unet.forward(wb, x[..., None].astype(np.float64))
    assert _rel(y, y64[..., 0]) < 1e-3


@pytest.mark.parametrize("n,h,w", [(1, 64, 96), (2, 24, 40)])
def test_fused_first_layer_is_identical(n, h, w):
    wts = unet.init_weights(3)
    net = ca.UnetNet(unet.flatten(wts))
    x = np.random.default_rng(7).standard_normal((n, h, w)).astype(np.float32)
    xd = _dev(x)
    y0 = torch.empty_like(xd)
    y1 = torch.empty_like(xd)
    net.forward(xd, y0, n, h, w)
    net.apply(xd, y1, n, h, w)
    np.testing.assert_array_equal(y0.cpu().numpy(), y1.cpu().numpy())


@pytest.mark.parametrize("n,h,w", [(2, 32, 64), (1, 24, 40)])
def test_backward_matches_oracle(n, h, w):
    print(f"backward n={n} h={h} w={w}", flush=True)
    wts = unet.init_weights(2)
    net = ca.UnetNet(unet.flatten(wts))
    rng = np.random.default_rng(n * h + w)
    x = rng.standard_normal((n, h, w)).astype(np.float32)
    gy = rng.standard_normal((n, h, w)).astype(np.float32)
    g = unet.unflatten(net.backward(_dev(x), _dev(gy), n, h, w))
    _, tape = unet.forward_device(_w64(wts), x[..., None])
    ref = unet.backward_device(_w64(wts), tape, gy[..., None].astype(np.float64))
    errs = {name: _rel(g[name], ref[name]) for name, _, _ in unet.LAYERS}
    print("gradient errors", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for name, err in errs.items():
        assert err < 2e-3, name


def _laplacian_pairs(n, h, w, seed):
    """Jacobi-scaled d = 1 p' operator (cell 0 pinned) and n error/residual pairs."""
    deg = np.full((h, w), 4.0)
    deg[0, :] -= 1
    deg[-1, :] -= 1
    deg[:, 0] -= 1
    deg[:, -1] -= 1
    deg[0, 0] = 1.0
    ce = np.zeros((h, w))
    cn = np.zeros((h, w))
    ce[:, :-1] = 1.0 / np.sqrt(deg[:, :-1] * deg[:, 1:])
    cn[:-1, :] = 1.0 / np.sqrt(deg[:-1, :] * deg[1:, :])
    ce[0, 0] = cn[0, 0] = 0.0
    rng = np.random.default_rng(seed)
    e = rng.standard_normal((n, h, w))
    e[:, 0, 0] = 0.0
    for _ in range(3):  # a little smoothing: errors of a Krylov solve are smooth-ish
        e = 0.5 * e + 0.125 * (np.roll(e, 1, 1) + np.roll(e, -1, 1) + np.roll(e, 1, 2) + np.roll(e, -1, 2))
        e[:, 0, 0] = 0.0
    r = e.copy()
    r[:, :, :-1] -= ce[:, :-1] * e[:, :, 1:]
    r[:, :, 1:] -= ce[:, :-1] * e[:, :, :-1]
    r[:, :-1, :] -= cn[:-1, :] * e[:, 1:, :]
    r[:, 1:, :] -= cn[:-1, :] * e[:, :-1, :]
    s = 1.0 / np.sqrt((r ** 2).mean(axis=(1, 2)))
    e *= s[:, None, None]
    r *= s[:, None, None]
    den = (e * r).sum(axis=(1, 2))
    return e, r, ce, cn, den


def test_loss_matches_oracle_and_training_reduces_it():
    n, h, w = 4, 32, 32
    e, r, ce, cn, den = _laplacian_pairs(n, h, w, 5)
    wts = unet.init_weights(4, scale=0.5)
    net = ca.UnetNet(unet.flatten(wts))
    xd, ed = _dev(r), _dev(e)
    ced, cnd = _dev(ce), _dev(cn)
    off = torch.zeros(n, dtype=torch.int64, device="cuda")
    dend = _dev(den)
    l0 = net.train_step(xd, ed, ced, cnd, off, dend, n, h, w, 0.0, update=False)
    yd = torch.empty_like(xd)
    net.forward(xd, yd, n, h, w)
    lref, _ = unet.anorm_loss(yd.cpu().numpy(), e, np.broadcast_to(ce, (n, h, w)),
                              np.broadcast_to(cn, (n, h, w)), den)
    assert l0 == pytest.approx(lref, rel=1e-4)
    for _ in range(40):
        net.train_step(xd, ed, ced, cnd, off, dend, n, h, w, 2e-3)
    l1 = net.train_step(xd, ed, ced, cnd, off, dend, n, h, w, 0.0, update=False)
    assert l1 < 0.7 * l0, (l0, l1)
    # the updated weights round-trip
    w1 = net.get_weights()
    assert np.isfinite(w1).all() and not np.array_equal(w1, unet.flatten(wts))


def _standalone(nx, ny, seed=20480001):
    """The d = 1 p' system of a cavity with a seeded mean-free RHS (b[0] = 0)."""
    disc = ca.GpuSimpleDiscretization.cavity(2, nx, ny)
    disc.upload("RESPONSE", np.ones(disc.n_cells))
    disc.upload("FLUX_STAR", np.zeros(disc.n_faces))
    disc.assemble_pressure_correction_resident(0)
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, disc.n_cells)
    b -= b.mean()
    b[0] = 0.0
    return disc, b


def test_learned_fcg_reaches_fp64_tolerance_and_falls_back():
    print("learned fcg 96x64", flush=True)
    disc, b = _standalone(96, 64)
    jac = ca.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi")
    ref = disc.solve_pressure_rhs(jac, b)
    # untrained weights: FCG (with its fallback if it stalls) still meets the FP64 target
    disc.set_learned_weights(unet.flatten(unet.init_weights(9, scale=0.3)))
    fcg = ca.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
    res = disc.solve_pressure_rhs(fcg, b)
    assert res.converged and res.true_residual <= 1e-8 * res.initial_residual * 1.0001
    np.testing.assert_allclose(res.solution, ref.solution, rtol=0,
                               atol=1e-6 * np.abs(ref.solution).max())
    # all-zero weights: M r = 0 off the pinned cell -> breakdown -> Jacobi-PCG finishes
    disc.set_learned_weights(np.zeros(ca.unet_param_count(), np.float32))
    res0 = disc.solve_pressure_rhs(fcg, b)
    assert res0.fallback_used and res0.converged
    assert res0.true_residual <= 1e-8 * res0.initial_residual * 1.0001
