"""Training of the learned p' preconditioner on the device.

The preconditioner is the linear 4-level U-Net of ``csrc/unet.hpp`` applied to
the Jacobi-scaled residual inside flexible CG. This module builds its training
set from the solver's own p' solves (``GpuSimpleDiscretization.harvest``),
initialises it as an additive multilevel operator and trains it with the
device kernels (``UnetNet.train_step``: BF16 MFMA forward and backward, A-norm
loss, Adam).

    python -m cfdapp_amd.learned --out cfdapp_amd/weights/pprime_unet.f32

Device memory comes from torch tensors (allocation only); every arithmetic
step runs in libcfdhip.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import (FIXED_GRADIENT, FIXED_VALUE, INLET, MOVING_WALL, OUTLET, WALL,
               GpuSimpleDiscretization, LinearSolverSettings, UnetNet, unet_param_count)

LAYERS = [("e0a", 1, 16), ("e0b", 16, 16), ("e1a", 16, 32), ("e1b", 32, 32),
          ("e2a", 32, 64), ("e2b", 64, 64), ("b3a", 64, 64), ("b3b", 64, 64),
          ("d2a", 128, 64), ("d2b", 64, 64), ("d1a", 96, 32), ("d1b", 32, 32),
          ("d0a", 48, 16), ("d0b", 16, 16), ("out", 16, 1)]


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# initialisation

def structured_init(seed=0, ratio=8.0, noise=0.0) -> np.ndarray:
    """An additive multilevel operator (BPX-like) as the starting point.

    Channel 0 carries each level's residual along the skip path, channel 1 a
    copy smoothed by the 3x3 binomial kernel K before it is averaged down, so
    restriction (average of K r) and prolongation (K of the nearest upsample)
    are transposes and the operator is symmetric positive definite:
        M r = sum_l ratio^l P_l R_l r.
    The other channels start at `noise` times a He-scaled random draw. Keep
    noise = 0 for a preconditioner: random weights make the operator
    non-symmetric, and flexible CG with one stored direction then loses most
    of its speed (measured: 17172 instead of 592 iterations at 2048^2 with
    noise 0.01); training needs a little noise to reach the unused channels.
    Returns the flat FP32 master vector (layer order, W[ky][kx][cin][cout]).
    """
    rng = np.random.default_rng(seed)
    k1 = np.array([1.0, 2.0, 1.0])
    K = np.outer(k1, k1) / 16.0
    w = {}
    for name, ci, co in LAYERS:
        w[name] = rng.standard_normal((3, 3, ci, co)) * math.sqrt(1.0 / (9 * ci)) * noise

    def ident(name, o, i, s=1.0):
        w[name][1, 1, i, o] += s

    def smooth(name, o, i, s=1.0):
        w[name][:, :, i, o] += s * K

    ident("e0a", 0, 0)
    ident("e0a", 1, 0)
    for b in ("e0b", "e1b", "e2b"):
        ident(b, 0, 0)
        smooth(b, 1, 1)
    for a in ("e1a", "e2a"):
        ident(a, 0, 1)
        ident(a, 1, 1)
    ident("b3a", 0, 1)
    smooth("b3b", 0, 0, ratio ** 3)
    smooth("d2a", 0, 0)
    ident("d2a", 0, 64, ratio ** 2)
    smooth("d1a", 0, 0)
    ident("d1a", 0, 64, ratio)
    smooth("d0a", 0, 0)
    ident("d0a", 0, 32, 1.0)
    for n in ("d2b", "d1b", "d0b", "out"):
        ident(n, 0, 0)
    flat = np.concatenate([w[name].astype(np.float32).ravel() for name, _, _ in LAYERS])
    assert flat.size == unet_param_count()
    return flat


class SymmetricMultilevel:
    """A symmetric parametrisation of the network: theta -> weights.

    The level-l term of M r = sum_l w_l P_l S_l R_l r uses a restriction
    R = average(K_j r) per level and a prolongation P = K_j(upsample) that is
    its transpose (K_j a 3 x 3 kernel symmetric under 90-degree rotations:
    centre, edge and corner weights), and a symmetric level kernel S_l. Every
    theta gives a symmetric operator, so flexible CG keeps the behaviour of
    preconditioned CG; training moves theta with the gradients the device's
    backward pass returns for the full weight set (dL/dtheta = J^T dL/dW,
    J by central differences of the construction, exact for its linear and
    bilinear terms).
    theta = [log w_0..3, K_0..2 (3 each), S_0..3 (3 each), log scale].
    """
    N_THETA = 4 + 9 + 12 + 1

    @staticmethod
    def initial(ratio=8.0):
        t = np.zeros(SymmetricMultilevel.N_THETA)
        t[0:4] = np.log([1.0, ratio, ratio ** 2, ratio ** 3])
        for j in range(3):                       # binomial kernels
            t[4 + 3 * j:7 + 3 * j] = [4 / 16, 2 / 16, 1 / 16]
        for j in range(3):                       # identity level kernels
            t[13 + 3 * j:16 + 3 * j] = [1.0, 0.0, 0.0]
        t[22:25] = [4 / 16, 2 / 16, 1 / 16]      # the coarsest level smooths
        t[25] = 0.0
        return t

    @staticmethod
    def kernel(c, e, k):
        return np.array([[k, e, k], [e, c, e], [k, e, k]])

    @staticmethod
    def weights(theta) -> np.ndarray:
        th = np.asarray(theta, np.float64)
        wl = np.exp(th[0:4])
        K = [SymmetricMultilevel.kernel(*th[4 + 3 * j:7 + 3 * j]) for j in range(3)]
        S = [SymmetricMultilevel.kernel(*th[13 + 3 * j:16 + 3 * j]) for j in range(4)]
        scale = np.exp(th[25])
        w = {name: np.zeros((3, 3, ci, co)) for name, ci, co in LAYERS}

        def ident(name, o, i, s=1.0):
            w[name][1, 1, i, o] += s

        ident("e0a", 0, 0)
        ident("e0a", 1, 0)
        for lvl, b in enumerate(("e0b", "e1b", "e2b")):
            ident(b, 0, 0)
            w[b][:, :, 1, 1] += K[lvl]
        for a in ("e1a", "e2a"):
            ident(a, 0, 1)
            ident(a, 1, 1)
        ident("b3a", 0, 1)
        w["b3b"][:, :, 0, 0] += wl[3] * S[3]
        w["d2a"][:, :, 0, 0] += K[2]
        w["d2a"][:, :, 64, 0] += wl[2] * S[2]
        w["d1a"][:, :, 0, 0] += K[1]
        w["d1a"][:, :, 64, 0] += wl[1] * S[1]
        w["d0a"][:, :, 0, 0] += K[0]
        w["d0a"][:, :, 32, 0] += wl[0] * S[0]
        for n in ("d2b", "d1b", "d0b"):
            ident(n, 0, 0)
        ident("out", 0, 0, scale)
        return np.concatenate([w[name].astype(np.float32).ravel() for name, _, _ in LAYERS])

    @staticmethod
    def _slot(name, ci, co):
        """Flat indices of the 3 x 3 kernel W[:, :, ci, co] of a layer, row-major (ky, kx)."""
        off = 0
        for nm, cin, cout in LAYERS:
            if nm == name:
                return np.array([off + ((ky * 3 + kx) * cin + ci) * cout + co
                                 for ky in range(3) for kx in range(3)])
            off += 9 * cin * cout
        raise KeyError(name)

    # centre / edge / corner positions of a 3 x 3 kernel (row-major)
    _MASKS = (np.array([0, 0, 0, 0, 1, 0, 0, 0, 0.]), np.array([0, 1, 0, 1, 0, 1, 0, 1, 0.]),
              np.array([1, 0, 1, 0, 0, 0, 1, 0, 1.]))

    @staticmethod
    def repair(theta, margin=0.01):
        """Keeps every level kernel S_l positive semi-definite: the symbol
        c + 2e (cos a + cos b) + 4k cos a cos b of [[k, e, k], [e, c, e], [k, e, k]]
        is bilinear in (cos a, cos b), so it is >= 0 on [-1, 1]^2 iff it is at
        the four corners: c >= 4|k| and c + 4k >= 4|e| (kept with a relative
        margin). With every S_l >= 0 and S_0 > 0 the network is SPD."""
        th = np.array(theta, np.float64)
        for lvl in range(4):
            c, e, k = th[13 + 3 * lvl:16 + 3 * lvl]
            c = max(c, 1e-6)
            k = float(np.clip(k, -(1 - margin) * c / 4, (1 - margin) * c / 4))
            lim = (1 - margin) * (c + 4 * k) / 4
            e = float(np.clip(e, -lim, lim))
            th[13 + 3 * lvl:16 + 3 * lvl] = (c, e, k)
        return th

    @staticmethod
    def theta_grad(theta, gw):
        """J^T gw: dL/dtheta from dL/dW (the construction is linear in the
        kernels and in the level weights, so J is exact)."""
        S_ = SymmetricMultilevel
        th = np.asarray(theta, np.float64)
        gw = np.asarray(gw, np.float64)
        wl = np.exp(th[0:4])
        skip = [("d0a", 32), ("d1a", 64), ("d2a", 64), ("b3b", 0)]
        g = np.zeros(len(th))
        for lvl in range(4):
            name, ci = skip[lvl]
            gs = gw[S_._slot(name, ci, 0)]
            Sk = S_.kernel(*th[13 + 3 * lvl:16 + 3 * lvl]).ravel()
            g[lvl] = wl[lvl] * float(gs @ Sk)                  # d / d log w_l
            for q in range(3):
                g[13 + 3 * lvl + q] = wl[lvl] * float(gs @ S_._MASKS[q])
        restr = [("e0b", 1, 1), ("e1b", 1, 1), ("e2b", 1, 1)]
        prol = [("d0a", 0, 0), ("d1a", 0, 0), ("d2a", 0, 0)]
        for j in range(3):
            gk = gw[S_._slot(*restr[j])] + gw[S_._slot(*prol[j])]
            for q in range(3):
                g[4 + 3 * j + q] = float(gk @ S_._MASKS[q])
        g[25] = np.exp(th[25]) * float(gw[S_._slot("out", 0, 0)][4])
        return g


class VCycle:
    """The network as a symmetric V(1,1) multigrid cycle (Neumann padding).

    Level l (H/2^l, l = 0..3) has a damped-Jacobi weight omega_l and a
    symmetric 3 x 3 operator A_l (centre, edge, corner); K_l (l = 0..2) is the
    transfer kernel between levels l and l + 1: restriction R = 4 pool(K .),
    prolongation P = K up(.) = R^T. With r_l the level's residual:
        encoder:  s_l = (I - omega_l A_l) r_l,  r_{l+1} = R s_l
        bottom:   x_3 = (I - omega_3 A_3)(omega_3 r_3 + c_3) + omega_3 r_3
        decoder:  x_l = (I - omega_l A_l)(omega_l r_l + P x_{l+1}) + omega_l r_l
    (one Jacobi sweep from zero, coarse correction, one sweep back), output
    y = scale x_0. c_3 is the coarse tail's correction (options "tail":
    the same cycle with fixed standard coefficients down to the coarsest
    grid, written into channel 2 of b3a) or 0. Channel use: encoder channel
    0 carries r_l along the skip path, channel 1 carries s_l (then K s_l,
    pooled); decoder channel 0 carries the correction, channel 1 r_l. Every
    kernel is mirror-symmetric in both axes and the cycle is symmetric positive
    semi-definite when omega_l (max of A_l's symbol) <= 2 (repair()).
    theta = [omega_0..3, A_0..3 (3 each), K_0..2 (3 each), log scale].
    """
    N_THETA = 4 + 12 + 9 + 1
    LAP = (1.0, -0.25, 0.0)
    BINOMIAL = (0.25, 0.125, 0.0625)
    # the tail's fixed coefficients (exact in BF16 and FP32)
    TAIL = {"levels": 16, "omega": 0.75, "a": LAP, "k": BINOMIAL}

    @staticmethod
    def options(tail=True):
        return {"padding": "neumann", "tail": dict(VCycle.TAIL) if tail else None}

    @staticmethod
    def initial(omega=0.75):
        t = np.zeros(VCycle.N_THETA)
        t[0:4] = omega
        for l in range(4):
            t[4 + 3 * l:7 + 3 * l] = VCycle.LAP
        for j in range(3):
            t[16 + 3 * j:19 + 3 * j] = VCycle.BINOMIAL
        t[25] = 0.0
        return t

    @staticmethod
    def kernel(c, e, k):
        return np.array([[k, e, k], [e, c, e], [k, e, k]])

    @staticmethod
    def weights(theta) -> np.ndarray:
        return VCycle.weights64(theta).astype(np.float32)

    @staticmethod
    def weights64(theta) -> np.ndarray:
        w = {name: np.zeros((3, 3, ci, co)) for name, ci, co in LAYERS}
        for (name, co, ci), k in VCycle.slots(theta).items():
            w[name][:, :, ci, co] += k
        return np.concatenate([w[name].ravel() for name, _, _ in LAYERS])

    @staticmethod
    def slots(theta) -> dict:
        """The non-zero 3 x 3 kernels: {(layer, cout, cin): kernel}."""
        th = np.asarray(theta, np.float64)
        om = th[0:4]
        A = [VCycle.kernel(*th[4 + 3 * l:7 + 3 * l]) for l in range(4)]
        K = [VCycle.kernel(*th[16 + 3 * j:19 + 3 * j]) for j in range(3)]
        scale = np.exp(th[25])
        d = VCycle.kernel(1.0, 0.0, 0.0)
        w = {}

        def put(name, co, ci, k):
            w[(name, co, ci)] = w.get((name, co, ci), 0.0) + k

        sweep = [d - om[l] * A[l] for l in range(4)]   # I - omega A
        put("e0a", 0, 0, d)
        put("e0a", 1, 0, sweep[0])
        for l, (a, b) in enumerate((("e1a", "e0b"), ("e2a", "e1b"), ("b3a", "e2b"))):
            put(b, 0, 0, d)               # r_l along the skip path
            put(b, 1, 1, K[l])            # K s_l, pooled into the next level
            put(a, 0, 1, 4.0 * d)         # r_{l+1} = 4 pool(K s_l)
            put(a, 1, 1, 4.0 * sweep[l + 1])
        put("b3b", 0, 0, om[3] * sweep[3] + om[3] * d)
        put("b3b", 0, 2, sweep[3])        # the tail's coarse correction
        for l, (a, b, ca) in enumerate((("d0a", "d0b", 32), ("d1a", "d1b", 64),
                                        ("d2a", "d2b", 64))):
            put(a, 0, 0, K[l])            # P x_{l+1}
            put(a, 0, ca, om[l] * d)      # + omega r_l
            put(a, 1, ca, d)              # r_l
            put(b, 0, 0, sweep[l])
            put(b, 0, 1, om[l] * d)
        put("out", 0, 0, scale * d)
        return w

    @staticmethod
    def repair(theta, margin=0.02):
        """0 < omega_l <= (2 - margin) / max(symbol of A_l): keeps the cycle PSD."""
        th = np.array(theta, np.float64)
        for l in range(4):
            c, e, k = th[4 + 3 * l:7 + 3 * l]
            top = max(c + 4 * e + 4 * k, c - 4 * e + 4 * k, c - 4 * k, 1e-6)
            th[l] = float(np.clip(th[l], 0.05, (2.0 - margin) / top))
        return th

    @staticmethod
    def theta_grad(theta, gw, h=1e-3):
        """J^T gw by central differences of the construction on its non-zero
        kernels (exact: they are at most quadratic in each parameter; the
        scale enters through exp and is differentiated exactly)."""
        th = np.asarray(theta, np.float64)
        gw = np.asarray(gw, np.float64)
        offs, o = {}, 0
        for name, ci, co in LAYERS:
            offs[name] = (o, ci, co)
            o += 9 * ci * co
        gs = {}
        for (name, co, ci) in VCycle.slots(th):
            base, cin, cout = offs[name]
            idx = base + (np.arange(9) * cin + ci) * cout + co
            gs[(name, co, ci)] = gw[idx].reshape(3, 3)
        g = np.zeros(len(th))
        for i in range(len(th) - 1):
            tp, tm = th.copy(), th.copy()
            tp[i] += h
            tm[i] -= h
            sp, sm = VCycle.slots(tp), VCycle.slots(tm)
            g[i] = sum(float(np.sum(gs[k] * (sp[k] - sm[k]))) for k in gs) / (2 * h)
        g[-1] = float(gs[("out", 0, 0)][1, 1]) * np.exp(th[25])
        return g


def train_symmetric(torch, bank, theta, steps, lr, seed=0, deadline=None, log_every=25,
                    eval_fn=None, eval_every=0, param=None, options=None):
    """Adam on theta with the device's loss gradients (param: the
    parametrisation, SymmetricMultilevel by default; options: how the network
    is applied). With eval_fn (theta's weights -> score, lower is better) the
    model is scored before training and every eval_every steps and the
    best-scoring theta is returned (model selection on held-out systems);
    returns (theta, losses, evals)."""
    P = param or SymmetricMultilevel
    net = UnetNet(P.weights(theta), options=options)
    rng = np.random.default_rng(seed)
    sizes = [k for k in bank.count if bank.count[k] > 0]
    probs = np.array([bank.count[k] * k[0] * k[0] for k in sizes], np.float64)
    probs /= probs.sum()
    m = np.zeros(len(theta))
    v = np.zeros(len(theta))
    b1, b2 = 0.9, 0.999
    losses, evals = [], []
    t0 = time.time()
    theta = np.array(theta, np.float64)
    best = (None, theta.copy(), 0)
    if eval_fn is not None:
        sc = eval_fn(P.weights(theta))
        evals.append((0, sc))
        best = (sc[0], theta.copy(), 0)
        _log(f"eval at 0: {sc}")
    for step in range(steps):
        if deadline is not None and time.time() > deadline:
            _log(f"train: time budget reached at step {step}")
            break
        key = sizes[int(rng.choice(len(sizes), p=probs))]
        n = key[0]
        b = min(BATCH[n], bank.count[key])
        idx = rng.choice(bank.count[key], size=b, replace=False)
        it = torch.from_numpy(idx).cuda()
        x = bank.r[key].index_select(0, it)
        e = bank.e[key].index_select(0, it)
        off = torch.from_numpy(bank.sys[key][idx] * n * n).cuda()
        den = torch.from_numpy(bank.den[key][idx]).cuda()
        if options and options.get("padding") == "neumann":
            net.set_options(dict(options, dirichlet_sides=list(key[1])))
        net.set_weights(P.weights(theta))
        loss, gw = net.loss_grad(x, e, bank.ce[key], bank.cn[key], off, den, b, n, n,
                                 cd=bank.cd[key])
        g = P.theta_grad(theta, gw)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        f = 0.5 * (1.0 + math.cos(math.pi * step / steps))
        theta -= lr * f * (m / (1 - b1 ** (step + 1))) / (np.sqrt(v / (1 - b2 ** (step + 1))) + 1e-8)
        theta = P.repair(theta)
        losses.append(loss)
        if step % log_every == 0 or step == steps - 1:
            _log(f"step {step}: loss {np.mean(losses[-log_every:]):.4f} ({n}^2 x{b}, "
                 f"{time.time() - t0:.0f} s) theta[0:4] {np.round(theta[0:4], 3).tolist()}")
        if eval_fn is not None and eval_every and ((step + 1) % eval_every == 0 or
                                                   step == steps - 1):
            sc = eval_fn(P.weights(theta))
            evals.append((step + 1, sc))
            _log(f"eval at {step + 1}: {sc}")
            if sc[0] < best[0]:
                best = (sc[0], theta.copy(), step + 1)
    if eval_fn is not None:
        _log(f"selected the model of step {best[2]} (score {best[0]:.4f})")
        return best[1], losses, evals
    return theta, losses, evals


TAIL_THETA0 = np.array([0.75, 1.0, -0.25, 0.0, 0.25, 0.125, 0.0625])


def tail_options(tail_theta, levels=16):
    """The coarse tail of the options from its 7 parameters: omega, the level
    operator (centre, edge, corner) and the transfer kernel (centre, edge,
    corner)."""
    t = np.asarray(tail_theta, np.float64)
    return {"padding": "neumann",
            "tail": {"levels": levels, "omega": float(t[0]), "a": [float(v) for v in t[1:4]],
                     "k": [float(v) for v in t[4:7]]}}


def repair_tail(tail_theta, margin=0.02):
    """0 < omega <= (2 - margin) / max symbol of the tail's level operator."""
    t = np.array(tail_theta, np.float64)
    c, e, k = t[1:4]
    top = max(c + 4 * e + 4 * k, c - 4 * e + 4 * k, c - 4 * k, 1e-6)
    t[0] = float(np.clip(t[0], 0.05, (2.0 - margin) / top))
    return t


def train_spectral(torch, bank, theta, tail_theta, steps, lr, k=6, seed=0, deadline=None,
                   log_every=25, eval_fn=None, eval_every=0, train_tail=True, fd_step=5e-3):
    """Adam on the V-cycle parameters (and the coarse tail's) with the
    multi-step spectral loss (UnetNet.spectral_loss_grad): the mean log
    convergence factor of the stationary iteration x <- x + N(b - A x) over
    k steps from the harvested errors, whose gradient follows the factor
    of the slowest error modes -- the ones that set the FCG iteration count --
    where the one-application A-norm loss follows the average error. The
    network weights' gradient comes from the device's backward pass (chain
    rule to theta on the host); the tail's 7 shared coefficients are not
    weights, and their gradient is the central difference of the same loss on
    the same batch. With eval_fn (weights, options -> score, lower is better)
    the model is scored before training and every eval_every steps, and the
    best-scoring parameters are returned (model selection on held-out
    systems): (theta, tail_theta, losses, evals)."""
    net = UnetNet(VCycle.weights(theta), options=tail_options(tail_theta))
    rng = np.random.default_rng(seed)
    sizes = [key for key in bank.count if bank.count[key] > 0]
    probs = np.array([bank.count[key] * key[0] * key[0] for key in sizes], np.float64)
    probs /= probs.sum()
    th = np.concatenate([np.asarray(theta, np.float64), np.asarray(tail_theta, np.float64)])
    nv = len(theta)
    m = np.zeros(len(th))
    v = np.zeros(len(th))
    b1, b2 = 0.9, 0.999
    losses, evals = [], []
    t0 = time.time()
    best = (None, th.copy(), 0)

    def opts_of(t, key):
        return dict(tail_options(t[nv:]), dirichlet_sides=list(key[1]))

    # the output scale does not change the FCG iterations but sets the
    # stationary iteration's convergence: start it at the scale that minimises
    # the loss on a few batches, so that the first steps do not spend the
    # other parameters on a diverging iteration
    batches = []
    for _ in range(4):
        key = sizes[int(rng.choice(len(sizes), p=probs))]
        n = key[0]
        b = min(BATCH[n], bank.count[key])
        idx = np.sort(rng.choice(bank.count[key], size=b, replace=False))
        batches.append((key, n, b, idx))

    def batch_loss(t):
        tot = 0.0
        for key, n, b, idx in batches:
            z = bank.e[key].index_select(0, torch.from_numpy(idx).cuda()).contiguous()
            off = torch.from_numpy(bank.sys[key][idx] * n * n).cuda()
            net.set_options(opts_of(t, key))
            net.set_weights(VCycle.weights(t[:nv]))
            tot += net.spectral_loss_grad(z, bank.ce[key], bank.cn[key], off, b, n, n, k,
                                          cd=bank.cd[key], grad=False)[0]
        return tot / len(batches)

    scales = np.linspace(-2.5, 0.5, 13)
    trial = []
    for sc in scales:
        t = th.copy()
        t[25] = sc
        trial.append(batch_loss(t))
    th[25] = float(scales[int(np.argmin(trial))])
    _log(f"output scale from the line search: exp({th[25]:.2f}), loss {min(trial):.4f}")
    if eval_fn is not None:
        sc = eval_fn(VCycle.weights(th[:nv]), tail_options(th[nv:]))
        evals.append((0, sc))
        best = (sc[0], th.copy(), 0)
        _log(f"eval at 0: {sc}")
    for step in range(steps):
        if deadline is not None and time.time() > deadline:
            _log(f"train: time budget reached at step {step}")
            break
        key = sizes[int(rng.choice(len(sizes), p=probs))]
        n = key[0]
        b = min(BATCH[n], bank.count[key])
        idx = np.sort(rng.choice(bank.count[key], size=b, replace=False))
        z = bank.e[key].index_select(0, torch.from_numpy(idx).cuda()).contiguous()
        off = torch.from_numpy(bank.sys[key][idx] * n * n).cuda()
        cd = bank.cd[key]
        net.set_options(opts_of(th, key))
        net.set_weights(VCycle.weights(th[:nv]))
        loss, gw = net.spectral_loss_grad(z, bank.ce[key], bank.cn[key], off, b, n, n, k, cd=cd)
        g = np.zeros(len(th))
        g[:nv] = VCycle.theta_grad(th[:nv], gw)
        if train_tail:
            for i in range(nv, len(th)):
                tp, tm = th.copy(), th.copy()
                tp[i] += fd_step
                tm[i] -= fd_step
                net.set_options(opts_of(tp, key))
                lp, _ = net.spectral_loss_grad(z, bank.ce[key], bank.cn[key], off, b, n, n, k,
                                               cd=cd, grad=False)
                net.set_options(opts_of(tm, key))
                lm, _ = net.spectral_loss_grad(z, bank.ce[key], bank.cn[key], off, b, n, n, k,
                                               cd=cd, grad=False)
                g[i] = (lp - lm) / (2 * fd_step)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        f = 0.5 * (1.0 + math.cos(math.pi * step / steps))
        th -= lr * f * (m / (1 - b1 ** (step + 1))) / (np.sqrt(v / (1 - b2 ** (step + 1))) + 1e-8)
        th[:nv] = VCycle.repair(th[:nv])
        th[nv:] = repair_tail(th[nv:])
        losses.append(loss)
        if step % log_every == 0 or step == steps - 1:
            _log(f"step {step}: spectral loss {np.mean(losses[-log_every:]):.4f} ({n}^2 x{b}, "
                 f"{time.time() - t0:.0f} s) omega {np.round(th[0:4], 3).tolist()} "
                 f"tail {np.round(th[nv:], 3).tolist()} |g| {np.abs(g[:nv]).max():.2e} "
                 f"{np.abs(g[nv:]).max():.2e}")
        if eval_fn is not None and eval_every and ((step + 1) % eval_every == 0 or
                                                   step == steps - 1):
            sc = eval_fn(VCycle.weights(th[:nv]), tail_options(th[nv:]))
            evals.append((step + 1, sc))
            _log(f"eval at {step + 1}: {sc}")
            if sc[0] < best[0]:
                best = (sc[0], th.copy(), step + 1)
    if eval_fn is not None:
        _log(f"selected the model of step {best[2]} (score {best[0]:.4f})")
        th = best[1]
    return th[:nv], th[nv:], losses, evals


class Validation:
    """Held-out p' systems for model selection (none is among the training
    cases, which use seed-drawn Re and speeds at 256^2-1024^2): a 256^2 cavity at
    Re 3200 after two SIMPLE iterations, a 512^2 channel at Re 1000 after one,
    and the d = 1 1024^2 cavity operator with a seeded mean-free RHS. The
    score is the mean log of the learned FCG iteration counts to 1e-8 (a
    solve that falls back or fails counts as 20000 iterations)."""

    def __init__(self):
        self.items = []
        d = make_case("cavity", 256, 3200.0, 1.0)
        assemble_pprime(d, 2)
        self.items.append(("cavity 256^2 Re 3200", d))
        d = make_case("channel", 512, 1000.0, 1.0)
        assemble_pprime(d, 1)
        self.items.append(("channel 512^2 Re 1000", d))
        d, b = standalone_system(1024, seed=1024)
        d.upload("PC_RHS", b)
        self.items.append(("d=1 1024^2", d))
        self.fcg = LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")

    def score(self, weights, options=None):
        its = []
        for _, d in self.items:
            d.set_learned_weights(weights, options=options or {"padding": "zero"})
            r = d.solve_pressure_rhs(self.fcg, None, download=False)
            its.append(r.iterations if r.status == 0 and not r.fallback_used else 20000)
        return float(np.mean(np.log(its))), its

    def close(self):
        for _, d in self.items:
            d.close()
        self.items = []


# ---------------------------------------------------------------------------
# cases and harvesting

def make_case(kind, n, re, speed, ny=None):
    """A 2D cavity (lid speed) or channel (inlet speed, left to right) of
    n x ny uniform cells with Re = speed L / nu, L the longer side. A square
    case is the unit square; a rectangle uses the cell size 2^-k that puts
    its longer side in (1/2, 1], so that every coordinate is exact (the
    structured path accepts only boundary faces exactly aligned with their
    owner cells)."""
    square = ny is None or ny == n
    ny = n if ny is None else ny
    h = 1.0 / n if square else 2.0 ** -math.ceil(math.log2(max(n, ny)))
    nu = speed * (max(n, ny) * h) / re
    d = GpuSimpleDiscretization()
    if kind == "cavity":
        vel = [(WALL, (0.0, 0.0, 0.0))] * 3 + [(MOVING_WALL, (speed, 0.0, 0.0))]
        pres = [(FIXED_GRADIENT, 0.0)] * 4
    else:
        vel = [(INLET, (speed, 0.0, 0.0)), (OUTLET, (0.0, 0.0, 0.0)),
               (WALL, (0.0, 0.0, 0.0)), (WALL, (0.0, 0.0, 0.0))]
        pres = [(FIXED_GRADIENT, 0.0), (FIXED_VALUE, 0.0), (FIXED_GRADIENT, 0.0),
                (FIXED_GRADIENT, 0.0)]
    ok, why = d.prepare(2, n, ny, lx=n * h, ly=ny * h, velocity=vel, pressure=pres,
                        viscosity=nu)
    if not ok:
        raise RuntimeError(why)
    return d


def assemble_pprime(d, warm_iterations, alpha_u=0.7):
    """Advances SIMPLE, then assembles the next iteration's p' system."""
    mom = LinearSolverSettings(1e-12, 1e-8, 2000, "BiCGSTAB")
    pres = LinearSolverSettings(0.0, 1e-8, 200000, "CG", "Jacobi", check_interval=64)
    if warm_iterations > 0:
        d.simple_run(warm_iterations, pressure=pres, momentum=mom)
    d._ctx.call("cfdhip_mass_flux")
    d.assemble_momentum_resident(alpha_u)
    for k in range(2):
        d.solve_momentum_resident(k, mom)
    d.compute_response_coefficients()
    d.compute_predicted_face_flux(True, alpha_u)
    d.assemble_pressure_correction_resident(0)


class PairBank:
    """Harvested pairs on the device, one group per key (n, sides): the grid
    size and, for the interior-scaled (Neumann-padded) network, the Dirichlet
    sides of the system's padding (its FixedValue pressure patches), so that
    every training batch shares one padding."""

    def __init__(self, torch, capacity, interior=False):
        self.torch = torch
        self.capacity = dict(capacity)  # n -> (max pairs, max systems) per group
        # interior: pairs for a network applied with the interior-equivalent
        # scaling (Neumann padding), see to_interior_scaling
        self.interior = interior
        self.r, self.e, self.ce, self.cn, self.den, self.sys = {}, {}, {}, {}, {}, {}
        self.cd = {}
        self.count, self.systems = {}, {}

    def _group(self, key):
        if key not in self.count:
            torch = self.torch
            n = key[0]
            pairs, systems = self.capacity[n]
            dev = dict(device="cuda", dtype=torch.float32)
            self.r[key] = torch.zeros((pairs, n, n), **dev)
            self.e[key] = torch.zeros((pairs, n, n), **dev)
            self.ce[key] = torch.zeros((systems, n, n), **dev)
            self.cn[key] = torch.zeros((systems, n, n), **dev)
            self.cd[key] = torch.ones((systems, n, n), **dev) if self.interior else None
            self.den[key] = np.zeros(pairs, np.float32)
            self.sys[key] = np.zeros(pairs, np.int64)
            self.count[key] = 0
            self.systems[key] = 0
        return key

    def key_of(self, d, n):
        if not self.interior:
            return (n, ())
        names = ("left", "right", "bottom", "top")
        return (n, tuple(names[p] for p in sorted(d._fixed_value) if p < 4))

    def harvest(self, d, n, stride, max_pairs, noise=0, seed=0):
        """Pairs from an FP64 Jacobi-PCG solve of d's assembled p' system, plus
        `noise` white-noise pairs on the same operator."""
        key = self._group(self.key_of(d, n))
        k0, s = self.count[key], self.systems[key]
        room = min(max_pairs + noise, self.capacity[n][0] - k0)
        if room <= 0 or s >= self.capacity[n][1]:
            return 0, None
        settings = LinearSolverSettings(0.0, 1e-10, 400000, "CG", "Jacobi", check_interval=64)
        got, den, res = d.harvest(settings, stride, max(0, room - noise), n, n, self.r[key][k0],
                                  self.e[key][k0], self.ce[key][s], self.cn[key][s])
        k1 = k0 + got
        wn = min(noise, self.capacity[n][0] - k1)
        if wn > 0:
            gn, dn = d.harvest_noise(seed, wn, n, n, self.r[key][k1], self.e[key][k1])
            den = np.concatenate([den, dn])
            got += gn
        self.den[key][k0:k0 + got] = den
        self.sys[key][k0:k0 + got] = s
        if self.interior:
            self.to_interior_scaling(d, key, k0, got, s)
        self.count[key] += got
        self.systems[key] += 1
        return got, res

    def to_interior_scaling(self, d, key, k0, got, s):
        """Rewrites system s and its pairs for a network applied as z = S N(S r)
        with S the library's scaling of a Neumann-padded network instead of
        D^-1/2: with F = (D S^2)^1/2 the network input is F r^, its target
        e^ / F, and the loss operator F A^ F (diagonal F^2, couplings
        F_i c_ij F_j); (e^, r^) is unchanged."""
        torch = self.torch
        f = torch.from_numpy(interior_factor(d).astype(np.float32)).cuda()
        self.r[key][k0:k0 + got] *= f
        self.e[key][k0:k0 + got] /= f
        self.ce[key][s][:, :-1] *= f[:, :-1] * f[:, 1:]
        self.cn[key][s][:-1, :] *= f[:-1, :] * f[1:, :]
        self.cd[key][s] = f * f

    def total(self):
        return sum(self.count.values())


def interior_factor(d, radius=4):
    """F = (D / D~)^1/2 on d's assembled 2D p' system (ny x nx): D the
    diagonal, D~ = S^-2 the library's scaling of a Neumann-padded network:
    the interior-equivalent diagonal (sum of internal-face couplings times 4
    / their count; D where a cell has none) geometric-mean filtered over a
    (2 radius + 1)^2 window with clamped indices; the pinned cell keeps D."""
    nx, ny = d.shape[0], d.shape[1]
    diag = d.download("PC_DIAG").reshape(ny, nx)
    coeff = d.download("PC_FACE_COEFF")
    fx = coeff[:(nx + 1) * ny].reshape(ny, nx + 1)
    fy = coeff[(nx + 1) * ny:].reshape(ny + 1, nx)
    s = np.zeros((ny, nx))
    m = np.zeros((ny, nx))
    s[:, 1:] += fx[:, 1:nx]
    m[:, 1:] += 1
    s[:, :-1] += fx[:, 1:nx]
    m[:, :-1] += 1
    s[1:, :] += fy[1:ny, :]
    m[1:, :] += 1
    s[:-1, :] += fy[1:ny, :]
    m[:-1, :] += 1
    dt = np.where((m > 0) & (s > 0), 4.0 * s / np.maximum(m, 1), diag)
    lg = np.log(dt)
    for axis in (1, 0):  # x then y, as the device does
        k = np.arange(-radius, radius + 1)
        size = lg.shape[axis]
        idx = np.clip(np.arange(size)[:, None] + k[None, :], 0, size - 1)
        lg = (np.take(lg, idx, axis=axis)).mean(axis=axis + 1)
    dt = np.exp(lg)
    if d.pin >= 0:
        dt.ravel()[d.pin] = diag.ravel()[d.pin]
    return np.sqrt(diag / dt)


def harvest_cases(torch, cases, capacity, seed=0, deadline=None, noise=0, interior=False):
    """Runs SIMPLE cases and harvests p' pairs; cases: list of (kind, n, re, speed, warm);
    `noise` white-noise pairs per case besides the solve's own."""
    bank = PairBank(torch, capacity, interior=interior)
    t0 = time.time()
    for i, (kind, n, re, speed, warm) in enumerate(cases):
        if deadline is not None and time.time() > deadline:
            _log(f"harvest: time budget reached after {i} cases")
            break
        d = make_case(kind, n, re, speed)
        try:
            assemble_pprime(d, warm)
            stride = max(4, n // 16)
            got, res = bank.harvest(d, n, stride, 200, noise=noise, seed=seed * 1000 + i)
        finally:
            d.close()
        it = res.iterations if res is not None else 0
        _log(f"harvest {i}: {kind} {n}^2 Re {re:.0f} U {speed:.2f} warm {warm}: "
             f"{it} it, {got} pairs (total {bank.total()}, {time.time() - t0:.0f} s)")
    return bank


def default_cases(seed, n_cases):
    """Cavities and channels at 256^2-1024^2, Re log-uniform in 100-5000,
    lid/inlet speed uniform in 0.5-2, 0-4 SIMPLE iterations before the harvest."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cases):
        kind = "cavity" if i % 2 == 0 else "channel"
        n = int(rng.choice([256, 512, 1024], p=[0.55, 0.33, 0.12]))
        re = float(math.exp(rng.uniform(math.log(100.0), math.log(5000.0))))
        speed = float(rng.uniform(0.5, 2.0))
        warm = int(rng.integers(0, 5))
        out.append((kind, n, re, speed, warm))
    return out


# ---------------------------------------------------------------------------
# training

BATCH = {256: 16, 512: 4, 1024: 1}


def train(torch, bank, weights, steps, lr, seed=0, deadline=None, log_every=50):
    net = UnetNet(weights)
    rng = np.random.default_rng(seed)
    sizes = [k for k in bank.count if bank.count[k] > 0]
    probs = np.array([bank.count[k] * k[0] * k[0] for k in sizes], np.float64)
    probs /= probs.sum()
    losses = []
    t0 = time.time()
    warm = max(1, steps // 20)
    for step in range(steps):
        if deadline is not None and time.time() > deadline:
            _log(f"train: time budget reached at step {step}")
            break
        key = sizes[int(rng.choice(len(sizes), p=probs))]
        n = key[0]
        b = min(BATCH[n], bank.count[key])
        idx = rng.choice(bank.count[key], size=b, replace=False)
        it = torch.from_numpy(idx).cuda()
        x = bank.r[key].index_select(0, it)
        e = bank.e[key].index_select(0, it)
        off = torch.from_numpy(bank.sys[key][idx] * n * n).cuda()
        den = torch.from_numpy(bank.den[key][idx]).cuda()
        # linear warm-up, cosine decay
        f = min(1.0, (step + 1) / warm) * 0.5 * (1.0 + math.cos(math.pi * step / steps))
        loss = net.train_step(x, e, bank.ce[key], bank.cn[key], off, den, b, n, n, lr * f)
        losses.append(loss)
        if step % log_every == 0 or step == steps - 1:
            recent = float(np.mean(losses[-log_every:]))
            _log(f"step {step}: loss {recent:.4f} (lr {lr * f:.2e}, {n}^2 x{b}, "
                 f"{time.time() - t0:.0f} s)")
    return net.get_weights(), losses


# ---------------------------------------------------------------------------
# evaluation

def standalone_system(n, seed=20480001):
    """The d = 1 p' system of an n x n cavity with a seeded mean-free RHS."""
    d = GpuSimpleDiscretization.cavity(2, n, n)
    d.upload("RESPONSE", np.ones(d.n_cells))
    d.upload("FLUX_STAR", np.zeros(d.n_faces))
    d.assemble_pressure_correction_resident(0)
    b = np.random.default_rng(seed).uniform(-1.0, 1.0, d.n_cells)
    b -= b.mean()
    b[0] = 0.0
    return d, b


def evaluate(weights, n, tol=1e-8, options=None):
    d, b = standalone_system(n)
    try:
        d.upload("PC_RHS", b)
        jac = LinearSolverSettings(0.0, tol, 400000, "CG", "Jacobi", check_interval=64)
        fcg = LinearSolverSettings(0.0, tol, 20000, "FCG", "Learned")
        d.set_learned_weights(weights, options=options or {"padding": "zero"})
        rj = d.solve_pressure_rhs(jac, None, download=False)
        rf = d.solve_pressure_rhs(fcg, None, download=False)
        return dict(n=n, jacobi_iterations=rj.iterations, jacobi_ms=rj.device_ms,
                    jacobi_status=rj.status_name,
                    jacobi_true_rel=rj.true_residual / rj.initial_residual,
                    fcg_iterations=rf.iterations, fcg_ms=rf.device_ms,
                    fcg_fallback=rf.fallback_used, fcg_status=rf.status_name,
                    fcg_true_rel=rf.true_residual / rf.initial_residual)
    finally:
        d.close()


def time_forward(n, passes, weights=None):
    """Wall time of the network's forward pass on a 1 x n x n input (ms)."""
    import torch
    net = UnetNet(structured_init(0, noise=0.0) if weights is None else weights)
    x = torch.randn(1, n, n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(2):
        net.forward(x, y, 1, n, n)
    t0 = time.perf_counter()
    for _ in range(passes):
        net.forward(x, y, 1, n, n)
    return 1e3 * (time.perf_counter() - t0) / passes


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--time-forward", type=int, default=0,
                    help="only time the forward pass on an n x n grid")
    ap.add_argument("--passes", type=int, default=8)
    ap.add_argument("--out", default=None)
    ap.add_argument("--cases", type=int, default=120)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--harvest-seconds", type=float, default=300)
    ap.add_argument("--train-seconds", type=float, default=420)
    ap.add_argument("--init", default=None, help="start from these weights instead")
    ap.add_argument("--free", action="store_true",
                    help="train every weight (zero padding; default: a symmetric parametrisation)")
    ap.add_argument("--param", choices=["vcycle", "multilevel"], default="vcycle",
                    help="vcycle: the V(1,1) cycle with Neumann padding and the coarse tail; "
                         "multilevel: the additive (BPX-like) family with zero padding")
    ap.add_argument("--noise", type=float, default=1e-3,
                    help="random weights in the unused channels of the initial network")
    ap.add_argument("--eval", type=int, nargs="*", default=[256, 1024])
    ap.add_argument("--noise-pairs", type=int, default=20,
                    help="white-noise pairs per harvested system")
    ap.add_argument("--checkpoints", type=int, default=10,
                    help="validation scores during training (model selection)")
    ap.add_argument("--summary", default=None)
    ap.add_argument("--loss", choices=["spectral", "anorm"], default="spectral",
                    help="spectral: the k-step convergence factor of the stationary iteration "
                         "(V-cycle only); anorm: the one-application A-norm error")
    ap.add_argument("--k", type=int, default=6, help="steps of the spectral loss")
    ap.add_argument("--fixed-tail", action="store_true",
                    help="keep the coarse tail's coefficients at their start")
    args = ap.parse_args(argv)
    if args.time_forward:
        ms = time_forward(args.time_forward, args.passes)
        print(json.dumps({"forward_ms": ms, "n": args.time_forward, "passes": args.passes}))
        return
    if not args.out:
        ap.error("--out is required for training")
    import torch
    cap = {256: (13000, 200), 512: (6500, 100), 1024: (1600, 30)}
    cases = default_cases(args.seed, args.cases)
    summary = dict(args=vars(args))
    vcycle = args.param == "vcycle" and not args.free
    param = VCycle if vcycle else SymmetricMultilevel
    opts = VCycle.options() if vcycle else {"padding": "zero"}
    if args.free:
        w0 = (np.fromfile(args.init, dtype="<f4") if args.init
              else structured_init(args.seed, noise=args.noise))
    else:
        theta0 = param.initial()
        w0 = param.weights(theta0)
    summary["eval_init"] = [evaluate(w0, n, options=opts) for n in args.eval]
    _log("init:", summary["eval_init"])
    bank = harvest_cases(torch, cases, cap, args.seed, time.time() + args.harvest_seconds,
                         noise=args.noise_pairs, interior=vcycle)
    summary["pairs"] = {f"{k[0]} {'+'.join(k[1]) or 'walls'}": int(c) for k, c in bank.count.items()}
    if args.free:
        w, losses = train(torch, bank, w0, args.steps, args.lr, args.seed,
                          time.time() + args.train_seconds)
    elif vcycle and args.loss == "spectral":
        val = Validation()
        try:
            theta, tail, losses, evals = train_spectral(
                torch, bank, theta0, TAIL_THETA0, args.steps, args.lr, k=args.k, seed=args.seed,
                deadline=time.time() + args.train_seconds,
                eval_fn=lambda wts, o: val.score(wts, o),
                eval_every=max(1, args.steps // args.checkpoints),
                train_tail=not args.fixed_tail)
        finally:
            val.close()
        summary["theta"] = theta.tolist()
        summary["tail_theta"] = tail.tolist()
        summary["evals"] = evals
        summary["parametrisation"] = "VCycle"
        summary["loss"] = f"spectral, {args.k} steps"
        opts = tail_options(tail)
        w = param.weights(theta)
    else:
        val = Validation()
        try:
            theta, losses, evals = train_symmetric(
                torch, bank, theta0, args.steps, args.lr, args.seed,
                time.time() + args.train_seconds,
                eval_fn=lambda wts: val.score(wts, opts),
                eval_every=max(1, args.steps // args.checkpoints), param=param, options=opts)
        finally:
            val.close()
        summary["theta"] = theta.tolist()
        summary["evals"] = evals
        summary["parametrisation"] = param.__name__
        w = param.weights(theta)
    summary["options"] = opts
    w.astype("<f4").tofile(args.out)
    summary["loss_first"] = float(np.mean(losses[:50])) if losses else None
    summary["loss_last"] = float(np.mean(losses[-50:])) if losses else None
    summary["steps"] = len(losses)
    summary["eval_trained"] = [evaluate(w, n, options=opts) for n in args.eval]
    _log("trained:", summary["eval_trained"])
    # the record beside the weights: how they are applied (load_learned_options)
    # and how they were made
    import os
    with open(args.summary or os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
