"""Parameter sets of the configured workloads, shared by bench.py and the tests so that both
build exactly the same boxes."""
from __future__ import annotations

import numpy as np

BATCHED_SIZES = [32] * 16 + [48] * 16 + [64] * 16


def batched48(precision: int = 32, seed: int = 2024, dt: float = 0.005):
    """The 48-box batched workload: 16 boxes each of 32^3, 48^3 and 64^3, each with its own
    viscosity nu ~ U[0.01, 0.05] and initial spectrum peak k0 ~ U[2, 6] (seeded draws, nu first),
    Kolmogorov forcing f = (sin 4y, 0, 0), fixed step dt.  Returns (BoxGroup parameter dicts,
    k0 per box)."""
    rng = np.random.default_rng(seed)
    nus = rng.uniform(0.01, 0.05, len(BATCHED_SIZES))
    k0s = rng.uniform(2.0, 6.0, len(BATCHED_SIZES))
    params = [dict(n=n, precision=precision, nu=float(nu), kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0,
                   dt=dt) for n, nu in zip(BATCHED_SIZES, nus)]
    return params, [float(k) for k in k0s]


def init_batched48(boxes, k0s, seed: int = 100):
    """Box i gets the seeded solenoidal field of (seed, box i, k0_i), energy 0.5."""
    for i, b in enumerate(boxes):
        b.init_random(seed, i, k0s[i], 0.5)
