"""Synthetic inputs of the measured workloads (shared by bench.py and the tests).

* ``splitmix_uniform``: the portable generator the workloads are drawn with,
  splitmix64 (Steele, Lea, Flood 2014) -> (x >> 11) * 2^-53, seeded with a
  decimal seed; element k of the stream is the (k+1)-th output.
* ``standalone_rhs``: the right-hand side of the stand-alone p' solve,
  b_i = 2 u_i - 1 from seed 20480001, its mean (summed serially in cell
  order) subtracted, b_0 = 0 (the pinned reference cell).
* ``batch_cases``: the 256 uneven 2D cases of the batched solve (seed 5005).
"""
from __future__ import annotations

import numpy as np

STANDALONE_SEED = 20480001
BATCH_SEED = 5005


def splitmix_uniform(seed: int, count: int) -> np.ndarray:
    """The first `count` uniforms in [0, 1) of the splitmix64 stream of `seed`."""
    golden = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + golden * idx
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def serial_sum(values: np.ndarray) -> float:
    """Left-to-right FP64 sum in index order (the reference's loop order)."""
    s = 0.0
    for chunk in np.array_split(values, max(1, len(values) // 65536)):
        for v in chunk.tolist():
            s += v
    return s


def standalone_rhs(n_cells: int, seed: int = STANDALONE_SEED) -> np.ndarray:
    """b_i = 2 u_i - 1, serial mean removed, b_0 = 0."""
    b = 2.0 * splitmix_uniform(seed, n_cells) - 1.0
    b -= serial_sum(b) / n_cells
    b[0] = 0.0
    return b


def batch_cases(seed: int = BATCH_SEED, count: int = 256):
    """(kind, nx, ny, Re) of `count` independent 2D cases: nx, ny uniform in
    64..512 with aspect ratio at most 4 (ny redrawn otherwise), cavity or
    channel with equal probability, Re log-uniform in 100-3200."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(rng.integers(64, 513))
        while True:
            ny = int(rng.integers(64, 513))
            if max(nx, ny) <= 4 * min(nx, ny):
                break
        kind = "cavity" if rng.random() < 0.5 else "channel"
        re = float(np.exp(rng.uniform(np.log(100.0), np.log(3200.0))))
        out.append((kind, nx, ny, re))
    return out


def apply_pprime_2d(nx: int, ny: int, pin: int, diag: np.ndarray, face_coeff: np.ndarray,
                    x: np.ndarray) -> np.ndarray:
    """y = A x for a 2D p' operator held as diagonal + one coupling per face
    (x-faces j-outer i = 0..nx, then y-faces j = 0..ny i-inner; boundary
    faces carry none): y_c = d_c x_c - sum over internal faces of D_f x_nb,
    and the pinned row y_pin = d_pin x_pin. A host check of residuals, not
    the device's summation order."""
    X = x.reshape(ny, nx)
    fx = face_coeff[:(nx + 1) * ny].reshape(ny, nx + 1)
    fy = face_coeff[(nx + 1) * ny:].reshape(ny + 1, nx)
    y = diag.reshape(ny, nx) * X
    y[:, 1:] -= fx[:, 1:nx] * X[:, :-1]
    y[:, :-1] -= fx[:, 1:nx] * X[:, 1:]
    y[1:, :] -= fy[1:ny, :] * X[:-1, :]
    y[:-1, :] -= fy[1:ny, :] * X[1:, :]
    y = y.reshape(-1)
    if pin >= 0:
        y[pin] = diag[pin] * x[pin]
    return y
