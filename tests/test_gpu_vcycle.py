"""GPU parity of the Neumann-padded network, its coarse tail, and the V-cycle
preconditioner inside flexible CG.

The device kernels are compared with the CPU oracle (oracle/unet.py, padding
"neumann", the tail of cfdapp_amd.learned.VCycle) with the device's rounding
points (forward_device / backward_device); the bounds are those of the
zero-padded network's tests: 1e-4 relative (L2) on activations and outputs,
2e-3 on every layer's weight gradient. The tail runs in FP32 on the device
and in FP64 in the oracle, well inside those bounds. The solve tests check
the FP64 true-residual tolerance and that the V-cycle network needs a small,
grid-independent number of FCG iterations.
"""
import numpy as np
import pytest

from cfdapp_amd.learned import VCycle
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


def _mirror(w):
    """Every kernel made mirror-symmetric in both axes."""
    return {k: 0.25 * (v + v[::-1] + v[:, ::-1] + v[::-1, ::-1]) for k, v in w.items()}


def _vcycle_weights(seed):
    th = VCycle.repair(VCycle.initial() + np.random.default_rng(seed).normal(0, 0.05, 26))
    w = unet.unflatten(VCycle.weights(th))
    # a little noise in the used and unused channels keeps every path busy
    rng = np.random.default_rng(seed + 1)
    noise = {k: rng.standard_normal(v.shape) * 0.02 for k, v in w.items()}
    return {k: (w[k] + _mirror(noise)[k]).astype(np.float32) for k in w}


@pytest.mark.parametrize("h,w,tail,sides", [(61, 45, True, ()), (129, 129, True, ()),
                                            (37, 21, False, ("right",)),
                                            (100, 60, True, ("right",))])
def test_odd_sided_apply_matches_oracle(h, w, tail, sides):
    """The inference path on images whose sides are not multiples of 8 (the
    tensors rounded up, the padding at each level's own edges, the ghost pool
    at odd edges) against the oracle's network on the image itself."""
    print(f"odd-sided apply h={h} w={w} tail={tail} sides={sides}", flush=True)
    opts = dict(VCycle.options(tail=tail), dirichlet_sides=list(sides))
    wts = _vcycle_weights(h + 2 * w)
    net = ca.UnetNet(unet.flatten(wts), options=opts)
    x = np.random.default_rng(h * w).standard_normal((1, h, w)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty_like(xd)
    net.apply(xd, yd, 1, h, w)
    y = yd.cpu().numpy()
    yb, tape = unet.forward_device(_w64(wts), x[..., None], opts=opts)
    err = _rel(y, yb[..., 0])
    print(f"output error {err:.1e}", flush=True)
    if err >= 1e-4:  # name the first layer that departs (live pixels of each tensor)
        h8, w8 = (h + 7) // 8 * 8, (w + 7) // 8 * 8
        for name in ca.UnetNet.TENSORS[1:17]:
            ref = tape[name]
            lv = 0 if ref.shape[1] > (h8 >> 1) else (1 if ref.shape[1] > (h8 >> 2) else
                                                     (2 if ref.shape[1] > (h8 >> 3) else 3))
            dev = net.tensor(name).reshape(h8 >> lv, w8 >> lv, -1)
            dev = dev[:ref.shape[1], :ref.shape[2], :ref.shape[3]]
            print(f"  {name}: {_rel(dev, ref[0]):.1e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("n,h,w,tail", [(2, 32, 64, True), (1, 64, 96, True),
                                        (1, 24, 40, False), (1, 128, 128, True)])
def test_neumann_forward_matches_oracle(n, h, w, tail):
    print(f"neumann forward n={n} h={h} w={w} tail={tail}", flush=True)
    opts = VCycle.options(tail=tail)
    wts = _vcycle_weights(h + w)
    net = ca.UnetNet(unet.flatten(wts), options=opts)
    x = np.random.default_rng(h * w).standard_normal((n, h, w)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty_like(xd)
    net.forward(xd, yd, n, h, w)
    y = yd.cpu().numpy()
    yb, tape = unet.forward_device(_w64(wts), x[..., None], opts=opts)
    errs = {name: _rel(net.tensor(name), tape[name].ravel()) for name in ca.UnetNet.TENSORS[:17]}
    print("activation errors", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert _rel(y, yb[..., 0]) < 1e-4
    for name, err in errs.items():
        assert err < 1e-4, name
    # the fused first layer (the solve path) gives the same bits
    y1 = torch.empty_like(xd)
    net.apply(xd, y1, n, h, w)
    np.testing.assert_array_equal(y1.cpu().numpy(), y)


def test_neumann_general_kernels_forward():
    """Neumann padding with arbitrary (not symmetric) kernels in every
    staging path, forward only."""
    wts = unet.init_weights(9)
    opts = {"padding": "neumann"}
    net = ca.UnetNet(unet.flatten(wts), options=opts)
    n, h, w = 1, 40, 72
    x = np.random.default_rng(3).standard_normal((n, h, w)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty_like(xd)
    net.forward(xd, yd, n, h, w)
    yb, _ = unet.forward_device(_w64(wts), x[..., None], opts=opts)
    assert _rel(yd.cpu().numpy(), yb[..., 0]) < 1e-4


@pytest.mark.parametrize("n,h,w", [(2, 32, 64), (1, 64, 64)])
def test_neumann_backward_matches_oracle(n, h, w):
    print(f"neumann backward n={n} h={h} w={w}", flush=True)
    opts = VCycle.options(tail=True)
    wts = _vcycle_weights(n * h + w)
    net = ca.UnetNet(unet.flatten(wts), options=opts)
    rng = np.random.default_rng(n * h + w)
    x = rng.standard_normal((n, h, w)).astype(np.float32)
    gy = rng.standard_normal((n, h, w)).astype(np.float32)
    g = unet.unflatten(net.backward(_dev(x), _dev(gy), n, h, w))
    _, tape = unet.forward_device(_w64(wts), x[..., None], opts=opts)
    ref = unet.backward_device(_w64(wts), tape, gy[..., None].astype(np.float64), opts=opts)
    errs = {name: _rel(g[name], ref[name]) for name, _, _ in unet.LAYERS}
    print("gradient errors", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for name, err in errs.items():
        assert err < 2e-3, name


def test_neumann_backward_refuses_asymmetric_kernels():
    net = ca.UnetNet(unet.flatten(unet.init_weights(1)), options={"padding": "neumann"})
    x = _dev(np.ones((1, 16, 16)))
    with pytest.raises(ca.CfdHipError, match="mirror-symmetric"):
        net.backward(x, x, 1, 16, 16)


def _fcg(d, weights, options):
    d.set_learned_weights(weights, options=options)
    s = ca.LinearSolverSettings(0.0, 1e-8, 5000, "FCG", "Learned")
    return d.solve_pressure_rhs(s, None, download=False)


@pytest.mark.parametrize("n", [64, 256])
def test_vcycle_fcg_on_the_standalone_operator(n):
    """d = 1 cavity operator, seeded mean-free RHS: the V-cycle network meets
    the FP64 tolerance in a grid-independent handful of iterations."""
    from cfdapp_amd.learned import standalone_system
    d, b = standalone_system(n)
    try:
        d.upload("PC_RHS", b)
        r = _fcg(d, VCycle.weights(VCycle.initial()), VCycle.options())
        print(f"{n}^2: {r.status_name} {r.iterations} it, true {r.true_residual:.2e}", flush=True)
        assert r.status == 0 and not r.fallback_used
        assert r.true_residual <= 1e-8 * np.linalg.norm(b) * (1 + 1e-9)
        assert r.iterations <= 25
        x = d.download("PPRIME")
        jac = d.solve_pressure_rhs(ca.LinearSolverSettings(0.0, 1e-10, 100000, "CG", "Jacobi"),
                                   None, download=False)
        assert jac.status == 0
        xj = d.download("PPRIME")
        assert np.abs(x - xj).max() <= 1e-6 * np.abs(xj).max()
    finally:
        d.close()


def test_vcycle_fcg_on_simple_systems():
    """p' systems of SIMPLE iterations (variable d): a cavity with the pinned
    cell and a channel with a FixedValue outlet."""
    from cfdapp_amd import learned
    for kind, n, ny, warm in (("cavity", 128, 128, 2), ("channel", 128, 64, 1)):
        d = learned.make_case(kind, n, 1000.0, 1.0, ny=ny)
        try:
            learned.assemble_pprime(d, warm)
            r = _fcg(d, VCycle.weights(VCycle.initial()), VCycle.options())
            jac = d.solve_pressure_rhs(ca.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi"),
                                       None, download=False)
            print(f"{kind} {n}x{ny}: FCG {r.iterations} it ({r.status_name}, fallback "
                  f"{r.fallback_used}), Jacobi-PCG {jac.iterations} it", flush=True)
            assert r.status == 0 and not r.fallback_used
            assert r.true_residual <= 1e-8 * r.initial_residual * (1 + 1e-9)
            # the untrained cycle's fixed stencil does not see the variable
            # response coefficients or the outlet's Dirichlet face; it still
            # needs several times fewer iterations (measured: 51 vs 765 on the
            # cavity, 53 vs 380 on the channel)
            assert r.iterations * 5 < jac.iterations
        finally:
            d.close()


def test_interior_scaled_pairs_are_consistent():
    """Pairs rewritten for the interior-equivalent scaling keep their A-norm:
    with y = 0 the loss of every pair is (e', A' e') / den = 1, with A' = F A^ F
    and e' = e^ / F; the device loss with the diagonal matches the oracle."""
    from cfdapp_amd import learned
    bank = learned.PairBank(torch, {64: (8, 1)}, interior=True)
    d = learned.make_case("channel", 64, 800.0, 1.0)
    try:
        learned.assemble_pprime(d, 1)
        got, _ = bank.harvest(d, 64, 8, 6, noise=2, seed=3)
    finally:
        d.close()
    assert got == 8
    key = (64, ("right",))  # the channel's FixedValue outlet
    assert list(bank.count) == [key]
    e = bank.e[key][:got].cpu().numpy().astype(np.float64)
    ce = np.broadcast_to(bank.ce[key][0].cpu().numpy(), e.shape)
    cn = np.broadcast_to(bank.cn[key][0].cpu().numpy(), e.shape)
    cd = np.broadcast_to(bank.cd[key][0].cpu().numpy(), e.shape)
    den = bank.den[key][:got].astype(np.float64)
    for k in range(got):
        lk, _ = unet.anorm_loss(np.zeros((1,) + e.shape[1:]), e[k:k + 1], ce[k:k + 1],
                                cn[k:k + 1], den[k:k + 1], cd=cd[k:k + 1])
        assert lk == pytest.approx(1.0, rel=2e-3), k
    # device loss with the diagonal against the oracle, V-cycle network output
    opts = dict(VCycle.options(), dirichlet_sides=["right"])
    w = VCycle.weights(VCycle.initial())
    net = ca.UnetNet(w, options=opts)
    off = torch.zeros(got, dtype=torch.int64, device="cuda")
    x = bank.r[key][:got].contiguous()
    loss, _ = net.loss_grad(x, bank.e[key][:got].contiguous(), bank.ce[key][0], bank.cn[key][0],
                            off, _dev(den), got, 64, 64, cd=bank.cd[key][0])
    yd = torch.empty_like(x)
    net.forward(x, yd, got, 64, 64)
    lref, _ = unet.anorm_loss(yd.cpu().numpy(), e, ce, cn, den, cd=cd)
    assert loss == pytest.approx(lref, rel=1e-4)


@pytest.mark.parametrize("sides", [["right"], ["left", "bottom"]])
def test_dirichlet_sides_forward_and_backward_match_oracle(sides):
    print(f"dirichlet sides {sides}", flush=True)
    opts = dict(VCycle.options(), dirichlet_sides=sides)
    wts = _vcycle_weights(31)
    net = ca.UnetNet(unet.flatten(wts), options=opts)
    n, h, w = 2, 64, 32
    rng = np.random.default_rng(len(sides))
    x = rng.standard_normal((n, h, w)).astype(np.float32)
    gy = rng.standard_normal((n, h, w)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty_like(xd)
    net.forward(xd, yd, n, h, w)
    yb, tape = unet.forward_device(_w64(wts), x[..., None], opts=opts)
    assert _rel(yd.cpu().numpy(), yb[..., 0]) < 1e-4
    y1 = torch.empty_like(xd)
    net.apply(xd, y1, n, h, w)
    np.testing.assert_array_equal(y1.cpu().numpy(), yd.cpu().numpy())
    g = unet.unflatten(net.backward(xd, _dev(gy), n, h, w))
    ref = unet.backward_device(_w64(wts), tape, gy[..., None].astype(np.float64), opts=opts)
    errs = {name: _rel(g[name], ref[name]) for name, _, _ in unet.LAYERS}
    print("gradient errors", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for name, err in errs.items():
        assert err < 2e-3, name


def test_channel_outlet_uses_a_dirichlet_side():
    """A channel's FixedValue outlet gets the odd reflection on the solver
    context: the V-cycle network then needs a few dozen FCG iterations."""
    from cfdapp_amd import learned
    d = learned.make_case("channel", 256, 1000.0, 1.0, ny=128)
    try:
        learned.assemble_pprime(d, 1)
        r = _fcg(d, VCycle.weights(VCycle.initial()), VCycle.options())
        jac = d.solve_pressure_rhs(ca.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi"),
                                   None, download=False)
        print(f"channel 256x128: FCG {r.iterations} it, Jacobi-PCG {jac.iterations} it", flush=True)
        assert r.status == 0 and not r.fallback_used
        assert r.true_residual <= 1e-8 * r.initial_residual * (1 + 1e-9)
        assert r.iterations <= 40 and 10 * r.iterations < jac.iterations
    finally:
        d.close()
