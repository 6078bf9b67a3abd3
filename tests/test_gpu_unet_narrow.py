"""The channel-compacted inference path of the learned preconditioner.

When the weights leave channels unused (the shipped weights: at most two
live channels per tensor), inference keeps only the channels some output
depends on and runs narrow FP32 kernels (cfdapp_amd/csrc/unet_narrow.hip)
instead of the BF16 MFMA layers over every channel. Both compute the same
linear map; they differ only in rounding (the MFMA layers see each
activation as a hi + lo BF16 pair, the narrow kernels in FP32), so the
narrow output must agree with the MFMA output (options dense = True) and
with the oracle's network (oracle/unet.py forward_device) within the
network tests' 1e-4 relative bound, and must not equal the MFMA output bit
for bit (which would mean the narrow path did not run). Covered: the
shipped weights with the coarse tail (two live channels: 2 stored per
pixel; at 1100 x 1100 the tail's first level runs pass by pass before the
one-workgroup cycle), a three-channel network (4 stored per pixel), zero and Neumann
padding, Dirichlet sides, odd image sides, a batch of two, and the FCG solve
with either path.
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


def _apply(weights, opts, x):
    n, h, w = x.shape
    net = ca.UnetNet(weights, options=opts)
    try:
        xd = _dev(x)
        yd = torch.empty_like(xd)
        net.apply(xd, yd, n, h, w)
        torch.cuda.synchronize()
        return yd.cpu().numpy()
    finally:
        net.close()


def _three_channel_weights(seed, mirror):
    """Random weights that read and write channels 0..2 only (4 stored per pixel)."""
    rng = np.random.default_rng(seed)
    w = {}
    for name, cin, cout in unet.LAYERS:
        v = np.zeros((3, 3, cin, cout))
        split = {"d2a": 64, "d1a": 64, "d0a": 32}.get(name)
        ins = list(range(min(3, cin))) + (list(range(split, split + 3)) if split else [])
        outs = list(range(min(3, cout)))
        v[np.ix_(range(3), range(3), ins, outs)] = rng.standard_normal((3, 3, len(ins), len(outs)))
        v *= 0.4
        if mirror:
            v = 0.25 * (v + v[::-1] + v[:, ::-1] + v[::-1, ::-1])
        w[name] = v
    return w


# 1100 x 1100: the coarse tail's first level (69 x 69) is above the LDS cycle's
# 4096 pixels and runs pass by pass over the chip before the cycle
SHIPPED = [((1, 64, 96), ()), ((2, 32, 64), ()), ((1, 61, 45), ()), ((1, 129, 129), ()),
           ((1, 100, 60), ("right",)), ((1, 37, 21), ("left", "bottom")), ((1, 1100, 1100), ())]


@pytest.mark.parametrize("shape,sides", SHIPPED)
def test_shipped_weights_narrow_matches_mfma_and_oracle(shape, sides):
    print("case: narrow, shipped weights", shape, sides, flush=True)
    wts = ca.load_learned_weights()
    opts = dict(ca.load_learned_options(), dirichlet_sides=list(sides))
    x = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
    y_n = _apply(wts, opts, x)
    y_d = _apply(wts, dict(opts, dense=True), x)
    w64 = {k: v.astype(np.float64) for k, v in unet.unflatten(wts).items()}
    y_o, _ = unet.forward_device(w64, x[..., None], opts=opts)
    # the exact map: BF16-rounded weights, FP64 activations
    y_x = unet.forward({k: unet.bf(v) for k, v in w64.items()}, x[..., None], opts=opts)
    e_o, e_d, e_x = _rel(y_n, y_o[..., 0]), _rel(y_n, y_d), _rel(y_n, y_x[..., 0])
    print(f"narrow vs oracle {e_o:.2e}, narrow vs MFMA {e_d:.2e}, MFMA vs oracle "
          f"{_rel(y_d, y_o[..., 0]):.2e}, narrow vs exact {e_x:.2e}", flush=True)
    assert e_o < 1e-4 and e_d < 1e-4
    # FP32 activations: the narrow path is within FP32 rounding of the exact map
    assert e_x < 5e-6
    assert not np.array_equal(y_n, y_d), "the narrow path did not run"


@pytest.mark.parametrize("padding,mirror", [("zero", False), ("neumann", True)])
def test_three_channel_network_on_the_wide_narrow_kernels(padding, mirror):
    print("case: narrow, three live channels,", padding, flush=True)
    w = _three_channel_weights(11, mirror)
    flat = unet.flatten({k: v.astype(np.float32) for k, v in w.items()})
    opts = {"padding": padding}
    x = np.random.default_rng(4).standard_normal((1, 72, 136)).astype(np.float32)
    y_n = _apply(flat, opts, x)
    y_d = _apply(flat, dict(opts, dense=True), x)
    y_o, _ = unet.forward_device(w, x[..., None], opts=opts)
    e_o, e_d = _rel(y_n, y_o[..., 0]), _rel(y_n, y_d)
    print(f"narrow vs oracle {e_o:.2e}, narrow vs MFMA {e_d:.2e}", flush=True)
    assert e_o < 1e-4 and e_d < 1e-4
    assert not np.array_equal(y_n, y_d)


def test_dense_weights_keep_the_mfma_path():
    """Weights with more than four live channels per tensor run the MFMA
    layers whatever dense says: the two options give the same bits."""
    flat = unet.flatten({k: v.astype(np.float32) for k, v in unet.init_weights(3).items()})
    x = np.random.default_rng(5).standard_normal((1, 64, 64)).astype(np.float32)
    y0 = _apply(flat, {"padding": "zero"}, x)
    y1 = _apply(flat, {"padding": "zero", "dense": True}, x)
    np.testing.assert_array_equal(y0, y1)


def test_2048_narrow_matches_mfma():
    print("case: narrow 2048 x 2048", flush=True)
    wts = ca.load_learned_weights()
    opts = ca.load_learned_options()
    x = np.random.default_rng(7).standard_normal((1, 2048, 2048)).astype(np.float32)
    y_n = _apply(wts, opts, x)
    y_d = _apply(wts, dict(opts, dense=True), x)
    e = _rel(y_n, y_d)
    print(f"narrow vs MFMA at 2048^2: {e:.2e}", flush=True)
    assert e < 1e-4


@pytest.mark.parametrize("n", [96, 512])
def test_fcg_with_either_path(n):
    """The same learned FCG solve with the narrow kernels and with the MFMA
    layers: both reach the FP64 target, in iteration counts within rounding of
    each other (the narrow path, FP32 activations, is the more accurate of
    the two: it needs at most one iteration more, and measured 26 vs 28 at
    512^2, 20 vs 20 at 96^2)."""
    print("case: narrow FCG", n, flush=True)
    from cfdapp_amd import learned
    its = {}
    for dense in (False, True):
        d = learned.make_case("cavity", n, ny=n, re=400.0, speed=1.0)
        try:
            learned.assemble_pprime(d, 2)
            d.set_learned_weights(options=dict(ca.load_learned_options(), dense=dense))
            ch = d.learned_channels()
            print("stored channels:", ch, flush=True)
            if dense:
                assert ch["e0"] == 16 and ch["b3a"] == 64
            else:  # two live channels everywhere but b3a (its own, the tail's source and target)
                assert ch["e0a"] == 0 and ch["e0"] == 2 and ch["d0"] == 2 and ch["b3a"] == 4
            r = d.solve_pressure_rhs(ca.LinearSolverSettings(0.0, 1e-8, 2000, "FCG", "Learned"),
                                     None, download=False)
            print(f"dense={dense}: {r.status_name} {r.iterations} it, {r.device_ms:.2f} ms, "
                  f"fallback {r.fallback_used}", flush=True)
            assert r.converged and not r.fallback_used
            its[dense] = r.iterations
        finally:
            d.close()
    assert its[False] <= its[True] + 1
    assert its[False] >= 0.8 * its[True] - 1
