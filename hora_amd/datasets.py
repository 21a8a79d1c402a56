"""Seeded synthetic workloads (numpy PCG64), shared by tests and bench.py.

  bf_normal      -- i.i.d. Normal(0, sigma) rows (the demo data of
                    examples/src/demo.rs:6-11), 10,000 x 64 base, 1,000 queries.
  bf_sparse_u8   -- integer values 0-255 stored as f32, each element 0 with
                    probability 0.5 else uniform in 1..255 (a synthetic
                    stand-in for fashion-mnist-784), 60,000 x 784 base.
  gaussian_mixture -- centres ~ Normal(0, I); each point picks a centre
                    uniformly and adds Normal(0, sigma^2) noise (isotropic,
                    or per-centre diagonal sigma_j ~ U(lo, hi) when
                    `aniso=(lo, hi)`).  The per-centre sigmas are drawn from
                    a generator seeded by the centres themselves, so points
                    drawn in separate calls from the same centres (a base
                    and its queries) share them.
  unit_normal    -- Normal(0, 1) rows L2-normalised.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def bf_normal(n: int, d: int, seed: int, sigma: float = 10.0) -> np.ndarray:
    return (_rng(seed).standard_normal((n, d), dtype=np.float32) * np.float32(sigma)).astype(
        np.float32)


def bf_sparse_u8(n: int, d: int, seed: int) -> np.ndarray:
    r = _rng(seed)
    vals = r.integers(1, 256, size=(n, d), dtype=np.int32)
    zero = r.random((n, d), dtype=np.float32) < 0.5
    vals[zero] = 0
    return vals.astype(np.float32)


def gaussian_mixture(n: int, d: int, n_centres: int, seed: int, sigma: float = 0.3,
                     aniso: tuple | None = None, centres: np.ndarray | None = None):
    r = _rng(seed)
    if centres is None:
        centres = r.standard_normal((n_centres, d), dtype=np.float32)
    which = r.integers(0, n_centres, size=n)
    noise = r.standard_normal((n, d), dtype=np.float32)
    if aniso is None:
        noise *= np.float32(sigma)
    else:
        sig = mixture_sigmas(centres, aniso)
        noise *= sig[which]
    x = centres[which] + noise
    return x.astype(np.float32), centres, which


def mixture_sigmas(centres: np.ndarray, aniso: tuple) -> np.ndarray:
    """Per-centre diagonal sigmas ~ U(lo, hi), a function of the centres."""
    import hashlib

    c = np.ascontiguousarray(centres, dtype=np.float32)
    seed = int.from_bytes(hashlib.sha256(c.tobytes()).digest()[:8], "little")
    return _rng(seed).uniform(aniso[0], aniso[1], size=c.shape).astype(np.float32)


def unit_normal(n: int, d: int, seed: int) -> np.ndarray:
    x = _rng(seed).standard_normal((n, d), dtype=np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True).astype(np.float32)
    return x.astype(np.float32)
