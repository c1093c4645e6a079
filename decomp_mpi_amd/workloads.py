"""Seeded synthetic workloads for the batched and serving forms.

* batched_shapes: independent 3D transforms, each axis drawn from
  {32, 48, 64, 96, 128} (expected 398,688 points per transform);
* serving_requests: requests with Poisson arrival times at rate lam, each
  axis log-uniform in [32, 256] rounded to a multiple of 16, complex float
  with probability 0.7 else complex double, direction forward or backward
  with equal probability.

Inputs of both are uniform[-1, 1) in real and imaginary parts.
"""
import numpy as np

BATCH_AXES = (32, 48, 64, 96, 128)


def batched_shapes(count=256, seed=4):
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(BATCH_AXES, 3)) for _ in range(count)]


def serving_requests(count, lam, seed=5):
    """[(arrival_s, shape, precision, direction)], arrivals from time 0."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / lam, count)
    t = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(count):
        axes = np.exp(rng.uniform(np.log(32), np.log(256), 3))
        shape = tuple(int(min(256, max(32, 16 * round(a / 16)))) for a in axes)
        precision = 0 if rng.uniform() < 0.7 else 1
        direction = -1 if rng.uniform() < 0.5 else 1
        out.append((float(t[i]), shape, precision, direction))
    return out


def uniform_complex(rng, shape, precision):
    x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return x.astype(np.complex128 if precision else np.complex64)
