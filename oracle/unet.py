"""CPU oracle of the learned p' preconditioner -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference (it is one of the additions
on the hot path), so this restatement is pinned by mathematics instead: its
backward pass against central finite differences of its own loss
(tests/test_unet_oracle.py), linearity and the adjoint identity of each layer,
and the reference-independent Green's-function check of the trained
preconditioner (FCG iteration counts). The HIP kernels (cfdapp_amd/csrc/unet.hip)
are checked against this module on the same weights and inputs.

Network (all convolutions 3x3, zero padding, no bias, no activation: the
preconditioner is a linear operator, z = M r, so flexible CG sees a fixed M):
  level 0 (H x W):      e0a = conv(x, 1->16);  e0 = conv(e0a, 16->16)
  level 1 (H/2):        e1a = conv(pool(e0), 16->32); e1 = conv(e1a, 32->32)
  level 2 (H/4):        e2a = conv(pool(e1), 32->64); e2 = conv(e2a, 64->64)
  level 3 (H/8):        b3a = conv(pool(e2), 64->64); b3 = conv(b3a, 64->64)
  decoder:              d2a = conv([up(b3), e2], 128->64); d2 = conv(d2a, 64->64)
                        d1a = conv([up(d2), e1], 96->32);  d1 = conv(d1a, 32->32)
                        d0a = conv([up(d1), e0], 48->16);  d0 = conv(d0a, 16->16)
  output:               y = conv(d0, 16->1)
pool = 2x2 average, up = nearest-neighbour x2, [a, b] = channel concatenation
(upsampled tensor first). Activations are NHWC; weights W[ky, kx, cin, cout].

Options (``opts`` dict, the device's cfdhip_unet_options): ``padding``
"zero" (default) or "neumann" (a tap outside the image reads the nearest pixel
inside: the adjoint is then the transposed convolution on the grid extended by
one ghost ring, folded back onto the boundary pixels), ``dirichlet_sides``
(with "neumann": the sides -- left, right, bottom, top; bottom is row 0 --
where the ghost is the negated nearest pixel, zero at the face), and ``tail`` (a
TailParams-like dict: levels, omega, a, k) -- a one-channel symmetric V(1,1)
multigrid cycle below the bottom level that restricts channel 1 of b3a and
writes the prolonged correction into its channel 2 (``tail_apply``).
"""
from __future__ import annotations

import numpy as np

CHANNELS = (16, 32, 64, 64)

# (name, cin, cout) in execution order
LAYERS = [
    ("e0a", 1, 16), ("e0b", 16, 16),
    ("e1a", 16, 32), ("e1b", 32, 32),
    ("e2a", 32, 64), ("e2b", 64, 64),
    ("b3a", 64, 64), ("b3b", 64, 64),
    ("d2a", 128, 64), ("d2b", 64, 64),
    ("d1a", 96, 32), ("d1b", 32, 32),
    ("d0a", 48, 16), ("d0b", 16, 16),
    ("out", 16, 1),
]


def init_weights(seed=0, scale=1.0):
    """He-style init, seeded (numpy default_rng)."""
    rng = np.random.default_rng(seed)
    w = {}
    for name, cin, cout in LAYERS:
        std = scale * np.sqrt(1.0 / (9.0 * cin))
        w[name] = (rng.standard_normal((3, 3, cin, cout)) * std).astype(np.float32)
    return w


def _signs(neumann):
    """Ghost-ring factors (bottom, top, left, right): -1 on a Dirichlet side."""
    sides = () if neumann is True else tuple(neumann)
    return tuple(-1.0 if s in sides else 1.0 for s in ("bottom", "top", "left", "right"))


def _pad1(x, neumann):
    """One-pixel ring around the H x W axes: zeros; or (neumann: True or the
    Dirichlet sides) the nearest pixel, negated on a Dirichlet side (a corner
    ghost takes both factors)."""
    x = np.asarray(x, dtype=np.float64)
    if not neumann:
        return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)), mode="edge")
    sb, st, sl, sr = _signs(neumann)
    xp[:, 0] *= sb
    xp[:, -1] *= st
    xp[:, :, 0] *= sl
    xp[:, :, -1] *= sr
    return xp


def _fold1(xp, neumann=True):
    """Adjoint of the ring: ghost pixels add onto the pixel they copy, with its factor."""
    sb, st, sl, sr = _signs(neumann)
    x = xp[:, 1:-1, 1:-1, :].copy()
    x[:, 0, :, :] += sb * xp[:, 0, 1:-1, :]
    x[:, -1, :, :] += st * xp[:, -1, 1:-1, :]
    x[:, :, 0, :] += sl * xp[:, 1:-1, 0, :]
    x[:, :, -1, :] += sr * xp[:, 1:-1, -1, :]
    x[:, 0, 0, :] += sb * sl * xp[:, 0, 0, :]
    x[:, 0, -1, :] += sb * sr * xp[:, 0, -1, :]
    x[:, -1, 0, :] += st * sl * xp[:, -1, 0, :]
    x[:, -1, -1, :] += st * sr * xp[:, -1, -1, :]
    return x


def conv3x3(x, w, neumann=False):
    """x: (N, H, W, Cin), w: (3, 3, Cin, Cout) -> (N, H, W, Cout); zero or
    Neumann padding."""
    n, h, wd, cin = x.shape
    xp = _pad1(x, neumann)
    out = np.zeros((n, h, wd, w.shape[3]), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            out += xp[:, ky:ky + h, kx:kx + wd, :] @ w[ky, kx].astype(np.float64)
    return out


def conv3x3_input_grad(g, w, neumann=False):
    """Adjoint of conv3x3 in x: dL/dx given dL/dout = g."""
    n, h, wd, cout = g.shape
    if neumann:  # the transposed conv on the ghost-extended grid, folded back
        dxp = np.zeros((n, h + 2, wd + 2, w.shape[2]), dtype=np.float64)
        g64 = np.asarray(g, dtype=np.float64)
        for ky in range(3):
            for kx in range(3):
                dxp[:, ky:ky + h, kx:kx + wd, :] += g64 @ w[ky, kx].T.astype(np.float64)
        return _fold1(dxp, neumann)
    gp = np.zeros((n, h + 2, wd + 2, cout), dtype=np.float64)
    gp[:, 1:-1, 1:-1, :] = g
    dx = np.zeros((n, h, wd, w.shape[2]), dtype=np.float64)
    # out[y, x] += in[y + ky - 1, x + kx - 1] w[ky, kx]  =>
    # din[y', x'] += g[y' - ky + 1, x' - kx + 1] w[ky, kx]^T
    for ky in range(3):
        for kx in range(3):
            dx += gp[:, 2 - ky:2 - ky + h, 2 - kx:2 - kx + wd, :] @ w[ky, kx].T.astype(np.float64)
    return dx


def conv3x3_weight_grad(x, g, neumann=False):
    """dL/dw given the conv input x and dL/dout = g."""
    n, h, wd, cin = x.shape
    xp = _pad1(x, neumann)
    dw = np.zeros((3, 3, cin, g.shape[3]), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            a = xp[:, ky:ky + h, kx:kx + wd, :].reshape(-1, cin)
            dw[ky, kx] = a.T @ g.reshape(-1, g.shape[3])
    return dw


def pool(x):
    """2x2 average to the rounded-up halves. On a side of odd length the
    coarse pixel that straddles the edge averages only the pixels inside
    (a zero row or column beyond the last one): a quarter of the transposed
    cropped upsampling ``up``, as everywhere inside, so the V-cycle network
    stays symmetric (the device's rule for sides that are not multiples of 8)."""
    x = np.asarray(x, dtype=np.float64)
    n, h, w, c = x.shape
    if h % 2 or w % 2:
        x = np.pad(x, ((0, 0), (0, h % 2), (0, w % 2), (0, 0)))
        n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_grad(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(x, like=None):
    """Nearest x2; cropped to the fine image `like` (odd sides)."""
    u = np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)
    return u if like is None else u[:, :like.shape[1], :like.shape[2]]


def up_grad(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def _neumann(opts):
    """The padding of the options: False (zero), True (Neumann) or the Dirichlet sides."""
    if not opts or opts.get("padding", "zero") != "neumann":
        return False
    return tuple(opts.get("dirichlet_sides") or ()) or True


def _stencil(v, c, e, k, neumann=True):
    """Symmetric 3 x 3 stencil (centre, edge, corner) with the ring of _pad1, v: (N, H, W)."""
    p = _pad1(np.asarray(v, np.float64)[..., None], neumann)[..., 0]
    h, w = v.shape[1:]
    edges = p[:, 0:h, 1:w + 1] + p[:, 2:h + 2, 1:w + 1] + p[:, 1:h + 1, 0:w] + p[:, 1:h + 1, 2:w + 2]
    corners = p[:, 0:h, 0:w] + p[:, 0:h, 2:w + 2] + p[:, 2:h + 2, 0:w] + p[:, 2:h + 2, 2:w + 2]
    return c * v + e * edges + k * corners


def tail_apply(s, tail, neumann=True):
    """The coarse tail on (N, H, W) images: K up(V(4 pool(K s))) with V the
    symmetric V(1,1) cycle of the device's TailParams (the network's padding;
    levels: the rounded-up halves of H x W while both sides are at least 2, at
    most tail["levels"] of them)."""
    s = np.asarray(s, dtype=np.float64)
    om, a, k = tail["omega"], tail["a"], tail["k"]
    st = lambda v, kern: _stencil(v, *kern, neumann)  # noqa: E731
    h, w = s.shape[1:]
    L = 0
    while L < tail["levels"] and h >= 2 and w >= 2:  # rounded-up halves
        h, w, L = (h + 1) // 2, (w + 1) // 2, L + 1
    if L == 0:
        return np.zeros_like(s)

    def restrict(v):
        """4 pool(K v): the sum over the fine pixels each coarse pixel covers."""
        kv = st(v, k)
        n, hh, ww = kv.shape
        kv = np.pad(kv, ((0, 0), (0, hh % 2), (0, ww % 2)))
        return kv.reshape(n, (hh + 1) // 2, 2, (ww + 1) // 2, 2).sum(axis=(2, 4))

    def prolong(e, like):
        """K up(e) on the fine grid of `like` (copies beyond an odd edge dropped)."""
        u = np.repeat(np.repeat(e, 2, axis=1), 2, axis=2)[:, :like.shape[1], :like.shape[2]]
        return st(u, k)

    r = [restrict(s)]
    for _ in range(L - 1):
        r.append(restrict(r[-1] - om * st(r[-1], a)))
    e = 2.0 * om * r[L - 1] - om * om * st(r[L - 1], a)
    for t in range(L - 2, -1, -1):
        x2 = om * r[t] + prolong(e, r[t])
        e = x2 - om * st(x2, a) + om * r[t]
    return prolong(e, s)


def _tail_forward(b3a, opts):
    tail = (opts or {}).get("tail")
    if not tail or tail.get("levels", 0) == 0:
        return b3a
    b3a = np.array(b3a, dtype=np.float64)
    b3a[..., 2] = tail_apply(b3a[..., 1], tail, _neumann(opts))
    return b3a


def _tail_backward(g, opts):
    tail = (opts or {}).get("tail")
    if not tail or tail.get("levels", 0) == 0:
        return g
    g = np.array(g, dtype=np.float64)
    g[..., 1] += tail_apply(g[..., 2], tail, _neumann(opts))
    g[..., 2] = 0.0
    return g


def forward(w, x, keep=False, opts=None):
    """x: (N, H, W, 1). Returns y (N, H, W, 1) [and the tape]. Sides not
    divisible by 8 coarsen to the rounded-up half per level (pool's ghost
    rule, up's crop); the padding reflects at each level's edges."""
    nm = _neumann(opts)

    def cv(a, name):
        return conv3x3(a, w[name], nm)

    t = {"x": x.astype(np.float64)}
    t["e0a"] = cv(t["x"], "e0a")
    t["e0"] = cv(t["e0a"], "e0b")
    t["p0"] = pool(t["e0"])
    t["e1a"] = cv(t["p0"], "e1a")
    t["e1"] = cv(t["e1a"], "e1b")
    t["p1"] = pool(t["e1"])
    t["e2a"] = cv(t["p1"], "e2a")
    t["e2"] = cv(t["e2a"], "e2b")
    t["p2"] = pool(t["e2"])
    t["b3a"] = _tail_forward(cv(t["p2"], "b3a"), opts)
    t["b3"] = cv(t["b3a"], "b3b")
    t["c2"] = np.concatenate([up(t["b3"], t["e2"]), t["e2"]], axis=3)
    t["d2a"] = cv(t["c2"], "d2a")
    t["d2"] = cv(t["d2a"], "d2b")
    t["c1"] = np.concatenate([up(t["d2"], t["e1"]), t["e1"]], axis=3)
    t["d1a"] = cv(t["c1"], "d1a")
    t["d1"] = cv(t["d1a"], "d1b")
    t["c0"] = np.concatenate([up(t["d1"], t["e0"]), t["e0"]], axis=3)
    t["d0a"] = cv(t["c0"], "d0a")
    t["d0"] = cv(t["d0a"], "d0b")
    t["y"] = cv(t["d0"], "out")
    return (t["y"], t) if keep else t["y"]


def backward(w, t, gy, opts=None):
    """Gradients of a scalar loss with dL/dy = gy: returns (dW dict, dL/dx)."""
    g = {}
    dw = {}
    nm = _neumann(opts)

    def conv_back(name, inp, gout):
        dw[name] = conv3x3_weight_grad(inp, gout, nm)
        return conv3x3_input_grad(gout, w[name], nm)

    gd0 = conv_back("out", t["d0"], gy)
    gd0a = conv_back("d0b", t["d0a"], gd0)
    gc0 = conv_back("d0a", t["c0"], gd0a)
    c = t["d1"].shape[3]
    g["d1"] = up_grad(gc0[..., :c])
    g["e0"] = gc0[..., c:]
    gd1a = conv_back("d1b", t["d1a"], g["d1"])
    gc1 = conv_back("d1a", t["c1"], gd1a)
    c = t["d2"].shape[3]
    g["d2"] = up_grad(gc1[..., :c])
    g["e1"] = gc1[..., c:]
    gd2a = conv_back("d2b", t["d2a"], g["d2"])
    gc2 = conv_back("d2a", t["c2"], gd2a)
    c = t["b3"].shape[3]
    g["b3"] = up_grad(gc2[..., :c])
    g["e2"] = gc2[..., c:]
    gb3a = _tail_backward(conv_back("b3b", t["b3a"], g["b3"]), opts)
    gp2 = conv_back("b3a", t["p2"], gb3a)
    g["e2"] = g["e2"] + pool_grad(gp2)
    ge2a = conv_back("e2b", t["e2a"], g["e2"])
    gp1 = conv_back("e2a", t["p1"], ge2a)
    g["e1"] = g["e1"] + pool_grad(gp1)
    ge1a = conv_back("e1b", t["e1a"], g["e1"])
    gp0 = conv_back("e1a", t["p0"], ge1a)
    g["e0"] = g["e0"] + pool_grad(gp0)
    ge0a = conv_back("e0b", t["e0a"], g["e0"])
    gx = conv_back("e0a", t["x"], ge0a)
    return dw, gx


# ---------------------------------------------------------------------------
# Device arithmetic emulation. The device keeps weights in BF16 (round to
# nearest even) and activations in FP32; a convolution's input enters the MFMA
# as hi + lo BF16 halves (split2), the weight gradients use BF16 operands, and
# everything accumulates in FP32. These variants round at exactly those points,
# so what remains against the device is the FP32 accumulation order.

def bf(a):
    """Round to BF16 (nearest even) and back to float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def split2(a):
    """The hi + lo BF16 pair an FP32 activation becomes in the MFMA."""
    a = f32(a)
    hi = bf(a)
    return hi + bf(a - hi)


def forward_device(w, x, opts=None):
    """Forward pass with device rounding: returns (y, tape)."""
    wb = {k: bf(v) for k, v in w.items()}
    t = {"x": np.asarray(x, dtype=np.float64)}
    nm = _neumann(opts)

    def conv(name, inp):
        return conv3x3(split2(inp), wb[name], nm)

    def conv_pool(name, inp):
        acc = conv(name, inp)
        return f32(acc), f32(pool(acc))

    t["e0a"] = f32(conv3x3(f32(t["x"]), wb["e0a"], nm))
    t["e0"], t["p0"] = conv_pool("e0b", t["e0a"])
    t["e1a"] = f32(conv("e1a", t["p0"]))
    t["e1"], t["p1"] = conv_pool("e1b", t["e1a"])
    t["e2a"] = f32(conv("e2a", t["p1"]))
    t["e2"], t["p2"] = conv_pool("e2b", t["e2a"])
    t["b3a"] = f32(_tail_forward(f32(conv("b3a", t["p2"])), opts))
    t["b3"] = f32(conv("b3b", t["b3a"]))
    t["c2"] = np.concatenate([up(t["b3"], t["e2"]), t["e2"]], axis=3)
    t["d2a"] = f32(conv("d2a", t["c2"]))
    t["d2"] = f32(conv("d2b", t["d2a"]))
    t["c1"] = np.concatenate([up(t["d2"], t["e1"]), t["e1"]], axis=3)
    t["d1a"] = f32(conv("d1a", t["c1"]))
    t["d1"] = f32(conv("d1b", t["d1a"]))
    t["c0"] = np.concatenate([up(t["d1"], t["e0"]), t["e0"]], axis=3)
    t["d0a"] = f32(conv("d0a", t["c0"]))
    t["d0"] = f32(conv("d0b", t["d0a"]))
    t["y"] = f32(conv3x3(t["d0"], wb["out"], nm))
    return t["y"], t


def backward_device(w, t, gy, opts=None):
    """Weight gradients with device rounding (FP32 gradient tensors, split
    inputs to the transposed convolutions, BF16 operands in the weight sums)."""
    wb = {k: bf(v) for k, v in w.items()}
    dw = {}
    gy = f32(gy)
    nm = _neumann(opts)

    def back(name, inp, g):
        dw[name] = conv3x3_weight_grad(bf(inp), bf(g), nm)
        return conv3x3_input_grad(split2(g), wb[name], nm)

    def split(gc, c):  # transposed concat layer: coarse part summed 2x2, skip part
        return f32(up_grad(gc[..., :c])), f32(gc[..., c:])

    dw["out"] = conv3x3_weight_grad(bf(t["d0"]), bf(gy), nm)
    gd0 = f32(conv3x3_input_grad(gy, wb["out"], nm))
    gd0a = f32(back("d0b", t["d0a"], gd0))
    gd1, ge0s = split(back("d0a", t["c0"], gd0a), t["d1"].shape[3])
    gd1a = f32(back("d1b", t["d1a"], gd1))
    gd2, ge1s = split(back("d1a", t["c1"], gd1a), t["d2"].shape[3])
    gd2a = f32(back("d2b", t["d2a"], gd2))
    gb3, ge2s = split(back("d2a", t["c2"], gd2a), t["b3"].shape[3])
    gb3a = f32(_tail_backward(f32(back("b3b", t["b3a"], gb3)), opts))
    gp2 = f32(back("b3a", t["p2"], gb3a))
    ge2 = f32(ge2s + pool_grad(gp2))
    ge2a = f32(back("e2b", t["e2a"], ge2))
    gp1 = f32(back("e2a", t["p1"], ge2a))
    ge1 = f32(ge1s + pool_grad(gp1))
    ge1a = f32(back("e1b", t["e1a"], ge1))
    gp0 = f32(back("e1a", t["p0"], ge1a))
    ge0 = f32(ge0s + pool_grad(gp0))
    ge0a = f32(back("e0b", t["e0a"], ge0))
    dw["e0a"] = conv3x3_weight_grad(bf(t["x"]), bf(ge0a), nm)
    return dw


def flatten(w):
    """Master-weight vector in layer order (the device layout)."""
    return np.concatenate([np.asarray(w[name], dtype=np.float32).ravel()
                           for name, _, _ in LAYERS])


def unflatten(v):
    w, o = {}, 0
    for name, cin, cout in LAYERS:
        k = 9 * cin * cout
        w[name] = np.asarray(v[o:o + k], dtype=np.float32).reshape(3, 3, cin, cout)
        o += k
    return w


def anorm_loss(y, e, ce, cn, den, cd=None):
    """Mean over samples of (d, A d) / den with d = y - e, A the operator of
    diagonal cd (None: unit diagonal) and couplings ce (to x + 1) and cn (to
    y + 1); returns (L, dL/dy). y, e, ce, cn, cd: (N, H, W); den: (N,)."""
    d = np.asarray(y, np.float64) - np.asarray(e, np.float64)
    ad = d.copy() if cd is None else np.asarray(cd, np.float64) * d
    ad[:, :, :-1] -= ce[:, :, :-1] * d[:, :, 1:]
    ad[:, :, 1:] -= ce[:, :, :-1] * d[:, :, :-1]
    ad[:, :-1, :] -= cn[:, :-1, :] * d[:, 1:, :]
    ad[:, 1:, :] -= cn[:, :-1, :] * d[:, :-1, :]
    n = d.shape[0]
    per = (d * ad).sum(axis=(1, 2)) / den
    return float(per.mean()), 2.0 * ad / (den[:, None, None] * n)


def operator_apply(v, ce, cn, cd=None):
    """B v per image for (N, H, W) arrays: diagonal cd (None: unit), couplings
    ce (to x + 1) and cn (to y + 1) -- the operator of anorm_loss."""
    v = np.asarray(v, np.float64)
    out = v.copy() if cd is None else np.asarray(cd, np.float64) * v
    out[:, :, :-1] -= ce[:, :, :-1] * v[:, :, 1:]
    out[:, :, 1:] -= ce[:, :, :-1] * v[:, :, :-1]
    out[:, :-1, :] -= cn[:, :-1, :] * v[:, 1:, :]
    out[:, 1:, :] -= cn[:, :-1, :] * v[:, :-1, :]
    return out


def spectral_loss(w, z, ce, cn, steps, cd=None, opts=None, grad=True, device=False):
    """Multi-step spectral loss of x <- x + N(b - B x): v_0 = z / |z|_B,
    v_j = v_{j-1} - N(B v_{j-1}), L = mean_s log(|v_steps|_B^2) / (2 steps);
    with grad, dL/dW by reverse accumulation through the steps (N^T = N, a
    self-adjoint network). z, ce, cn, cd: (N, H, W). device: the network's
    passes with the device's rounding points (forward_device,
    backward_device). Returns (L, dW dict or None)."""
    z = np.asarray(z, np.float64)
    n = z.shape[0]
    if device:
        def net(x):
            return forward_device(w, x[..., None], opts=opts)[0][..., 0]

        def grad_of(x, g):
            _, tape = forward_device(w, x[..., None], opts=opts)
            return backward_device(w, tape, g[..., None], opts=opts)
    else:
        def net(x):
            return forward(w, x[..., None], opts=opts)[..., 0]

        def grad_of(x, g):
            _, tape = forward(w, x[..., None], keep=True, opts=opts)
            return backward(w, tape, g[..., None], opts=opts)[0]
    bz = operator_apply(z, ce, cn, cd)
    nz = np.sqrt((z * bz).sum(axis=(1, 2)))[:, None, None]
    v, bv = [z / nz], [bz / nz]
    for _ in range(steps):
        v.append(v[-1] - net(bv[-1]))
        bv.append(operator_apply(v[-1], ce, cn, cd))
    raw = (v[-1] * bv[-1]).sum(axis=(1, 2))
    loss = float(np.mean(np.log(raw)) / (2 * steps))
    if not grad:
        return loss, None
    g = bv[-1] / (steps * n * raw[:, None, None])
    dw = {name: np.zeros((3, 3, ci, co)) for name, ci, co in LAYERS}
    for j in range(steps, 0, -1):
        gj = grad_of(bv[j - 1], g)
        for name in dw:
            dw[name] -= gj[name]
        if j > 1:
            g = g - operator_apply(net(g), ce, cn, cd)
    return loss, dw


def mse_loss(y, target):
    """L = mean((y - target)^2); returns (L, dL/dy)."""
    d = y - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def laplacian_apply(z, pin=True):
    """The d = 1 p' operator on an (H, W) grid (Neumann faces carry no
    coupling, cell 0 pinned to the identity row), as a dense reference."""
    h, w = z.shape
    y = np.zeros_like(z, dtype=np.float64)
    deg = np.zeros_like(z, dtype=np.float64)
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        ys, ye = max(0, -dy), h - max(0, dy)
        xs, xe = max(0, -dx), w - max(0, dx)
        y[ys:ye, xs:xe] -= z[ys + dy:ye + dy, xs + dx:xe + dx]
        deg[ys:ye, xs:xe] += 1.0
    y += deg * z
    if pin:
        y[0, 0] = z[0, 0]
    return y
