"""GPU parity of the multi-step spectral training loss
(cfdhip_unet_spectral_loss_grad) against the CPU oracle (oracle/unet.py
spectral_loss, pinned by tests/test_spectral_oracle.py): the loss and its
gradient in the V-cycle parameters on Neumann-padded images with a Dirichlet
side and the coarse tail. The start parameters are exact in BF16, so the
device and the oracle apply the same network. The loss agrees with the exact
oracle to 1e-4 (FP32 activations split hi/lo at the MFMA against FP64). The
device's weight gradients take BF16 operands (one rounding of each activation
and gradient per product), so the parameter gradient is compared with the
oracle's device-rounding variant (forward_device / backward_device: the same
rounding points, FP64 sums) to 2 % of its largest component (a BF16 operand
that lands on the other side of a rounding boundary differs by one BF16 unit,
2^-8 relative, in either direction), and with the exact oracle to 8 % (the
rounding of a sum with cancellation across the steps); a wrong term in the
reverse accumulation moves components by their own size. A third check is
the device's own gradient against central differences of its own loss.
"""
import numpy as np
import pytest

from cfdapp_amd.learned import TAIL_THETA0, VCycle, tail_options
from oracle import unet as U

torch = pytest.importorskip("torch")
import cfdapp_amd as ca  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    torch.cuda.synchronize()
    return t


def _w64(theta):
    v, w, o = VCycle.weights64(theta), {}, 0
    for name, cin, cout in U.LAYERS:
        w[name] = v[o:o + 9 * cin * cout].reshape(3, 3, cin, cout)
        o += 9 * cin * cout
    return w


def _problem(n, h, w, seed):
    rng = np.random.default_rng(seed)
    ce = np.zeros((n, h, w))
    cn = np.zeros((n, h, w))
    ce[:, :, :-1] = rng.uniform(0.5, 1.5, (n, h, w - 1))
    cn[:, :-1, :] = rng.uniform(0.5, 1.5, (n, h - 1, w))
    cd = np.zeros((n, h, w))
    cd[:, :, :-1] += ce[:, :, :-1]
    cd[:, :, 1:] += ce[:, :, :-1]
    cd[:, :-1, :] += cn[:, :-1, :]
    cd[:, 1:, :] += cn[:, :-1, :]
    cd += 0.05
    z = rng.standard_normal((n, h, w))
    # float32 operators, as the training bank holds them
    return [a.astype(np.float32).astype(np.float64) for a in (z, ce, cn, cd)]


def _device_loss(net, z, ce, cn, cd, n, h, w, k, grad=True):
    off = torch.arange(n, dtype=torch.int64, device="cuda") * (h * w)
    return net.spectral_loss_grad(_dev(z), _dev(ce), _dev(cn), off, n, h, w, k, cd=_dev(cd),
                                  grad=grad)


def test_spectral_loss_and_gradient_match_oracle():
    n, h, w, k = 2, 32, 32, 3
    print(f"spectral loss parity n={n} h={h} w={w} k={k}", flush=True)
    z, ce, cn, cd = _problem(n, h, w, 5)
    th = VCycle.initial()
    opts = dict(tail_options(TAIL_THETA0), dirichlet_sides=["right"])
    net = ca.UnetNet(VCycle.weights(th), options=opts)
    try:
        loss, gw = _device_loss(net, z, ce, cn, cd, n, h, w, k)
    finally:
        net.close()
    ref, dw = U.spectral_loss(_w64(th), z, ce, cn, k, cd=cd, opts=opts)
    assert abs(loss - ref) < 1e-4, (loss, ref)
    g = VCycle.theta_grad(th, gw)
    gref = VCycle.theta_grad(th, U.flatten(dw))
    err = np.abs(g - gref).max()
    assert err <= 0.08 * np.abs(gref).max(), (err, np.abs(gref).max())
    _, dwd = U.spectral_loss(_w64(th), z, ce, cn, k, cd=cd, opts=opts, device=True)
    gdev = VCycle.theta_grad(th, U.flatten(dwd))
    err = np.abs(g - gdev).max()
    print(f"theta gradient: device vs rounding model {err:.3e}, vs exact "
          f"{np.abs(g - gref).max():.3e}, largest {np.abs(gref).max():.3e}", flush=True)
    assert err <= 0.02 * np.abs(gdev).max(), (err, np.abs(gdev).max())


def test_spectral_gradient_matches_its_own_differences():
    n, h, w, k = 2, 64, 64, 4
    print(f"spectral loss differences n={n} h={h} w={w} k={k}", flush=True)
    z, ce, cn, cd = _problem(n, h, w, 6)
    # a perturbed cycle; its weights are rounded to BF16 on the device, so the
    # difference quotient is taken on theta steps large next to that rounding
    th = VCycle.repair(VCycle.initial() + np.random.default_rng(2).uniform(-0.05, 0.05, 26))
    opts = dict(tail_options(TAIL_THETA0))
    net = ca.UnetNet(VCycle.weights(th), options=opts)
    try:
        loss, gw = _device_loss(net, z, ce, cn, cd, n, h, w, k)
        g = VCycle.theta_grad(th, gw)
        hs = 2e-2
        for i in (0, 2, 4, 16, 25):
            tp, tm = th.copy(), th.copy()
            tp[i] += hs
            tm[i] -= hs
            net.set_weights(VCycle.weights(tp))
            lp, _ = _device_loss(net, z, ce, cn, cd, n, h, w, k, grad=False)
            net.set_weights(VCycle.weights(tm))
            lm, _ = _device_loss(net, z, ce, cn, cd, n, h, w, k, grad=False)
            fd = (lp - lm) / (2 * hs)
            assert abs(fd - g[i]) <= 0.02 * np.abs(g).max() + 0.05 * abs(fd), (i, fd, g[i])
    finally:
        net.close()


def test_spectral_loss_rejects_a_zero_error_image():
    n, h, w = 1, 16, 16
    z, ce, cn, cd = _problem(n, h, w, 7)
    net = ca.UnetNet(VCycle.weights(VCycle.initial()), options=tail_options(TAIL_THETA0))
    try:
        with pytest.raises(ca.CfdHipError):
            _device_loss(net, np.zeros_like(z), ce, cn, cd, n, h, w, 2)
    finally:
        net.close()
