"""The symmetric parametrisation used to train the p' preconditioner.

Checks on the CPU that theta -> weights gives the structured multilevel
start at the initial theta, that every theta gives a symmetric linear
operator (the property flexible CG relies on), and that the chain rule
dL/dtheta = J^T dL/dW matches central differences.
"""
import numpy as np

from cfdapp_amd.learned import SymmetricMultilevel as S, structured_init
from oracle import unet as U


def test_initial_theta_is_structured_start():
    assert np.array_equal(S.weights(S.initial()), structured_init(ratio=8.0, noise=0.0))


def test_theta_grad_matches_central_differences():
    rng = np.random.default_rng(1)
    th = S.initial() + rng.normal(0.0, 0.05, S.N_THETA)
    gw = rng.normal(size=S.weights(th).size)
    g = S.theta_grad(th, gw)
    fd = np.zeros_like(g)
    for i in range(len(th)):
        a, b = th.copy(), th.copy()
        a[i] += 1e-5
        b[i] -= 1e-5
        fd[i] = (gw @ S.weights(a).astype(np.float64) - gw @ S.weights(b).astype(np.float64)) / 2e-5
    assert np.max(np.abs(g - fd)) <= 2e-3 * np.max(np.abs(fd))


def test_every_theta_gives_a_symmetric_operator():
    rng = np.random.default_rng(3)
    th = S.initial() + rng.normal(0.0, 0.2, S.N_THETA)
    params = U.unflatten(S.weights(th).astype(np.float64))
    h = w = 16
    x = rng.normal(size=(1, h, w, 1))
    y = rng.normal(size=(1, h, w, 1))
    nx = U.forward(params, x)
    ny = U.forward(params, y)
    # (y, N x) == (N y, x) for a symmetric N
    assert abs(np.sum(y * nx) - np.sum(ny * x)) <= 1e-9 * np.sqrt(np.sum(nx * nx) * np.sum(y * y))


def test_repair_keeps_level_kernels_positive_semidefinite():
    """After repair every level kernel has a nonnegative symbol on the whole
    frequency square, and the network is positive definite on random inputs."""
    rng = np.random.default_rng(4)
    a = np.linspace(0.0, np.pi, 33)
    X, Y = np.meshgrid(np.cos(a), np.cos(a))
    for _ in range(50):
        th = S.repair(S.initial() + rng.normal(0.0, 0.5, S.N_THETA))
        for lvl in range(4):
            c, e, k = th[13 + 3 * lvl:16 + 3 * lvl]
            assert np.min(c + 2 * e * (X + Y) + 4 * k * X * Y) >= -1e-12
    params = U.unflatten(S.weights(th).astype(np.float64))
    for _ in range(5):
        x = rng.normal(size=(1, 16, 16, 1))
        assert np.sum(x * U.forward(params, x)) > 0.0
    # the multilevel start is already feasible (up to the margin)
    assert np.max(np.abs(S.repair(S.initial()) - S.initial())) < 5e-3
