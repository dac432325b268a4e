"""Seeded synthetic workloads.

The reference ships no datasets usable offline (its loader downloads
ann-benchmarks HDF5 files, python/ann_dataset/dataset.py:28-38), so every
workload here is generated from fixed seeds with numpy's PCG64 generator,
which is reproducible across machines for a given numpy version.

Seeds: base 1234, queries 5678, per-query ef/k 91011, arrivals 1213.
"""

from __future__ import annotations

import numpy as np

BASE_SEED = 1234
QUERY_SEED = 5678
EFK_SEED = 91011
ARRIVAL_SEED = 1213


def gaussian(n: int, d: int, seed: int) -> np.ndarray:
    """i.i.d. N(0,1) rows (the 10k x 128 parity workload)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, d), dtype=np.float32)


def _mixture_centers(n_centers: int, d: int, low: float, high: float) -> np.ndarray:
    rng = np.random.default_rng(BASE_SEED + 7)
    return rng.uniform(low, high, size=(n_centers, d)).astype(np.float32)


def uint8_mixture(n: int, d: int, seed: int, n_centers: int = 1024, sigma: float = 20.0,
                  chunk: int = 1 << 17) -> np.ndarray:
    """Mixture of isotropic Gaussians, centers U[0,255]^d, clipped to [0,255]
    and rounded to integers (the 1M x 128 SQ4U+rerank workload)."""
    centers = _mixture_centers(n_centers, d, 0.0, 255.0)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, n_centers, size=e - s)
        x = out[s:e]
        rng.standard_normal(out=x, dtype=np.float32)
        x *= np.float32(sigma)
        x += centers[lab]
        np.clip(x, 0.0, 255.0, out=x)
        np.rint(x, out=x)
    return out


def normal_mixture(n: int, d: int, seed: int, n_centers: int = 4096, center_std: float = 2.0,
                   sigma: float = 1.0, chunk: int = 1 << 17) -> np.ndarray:
    """Mixture with centers N(0, center_std^2 I) and per-component sigma
    (the 10M x 128 SQ2U workload)."""
    crng = np.random.default_rng(BASE_SEED + 11)
    centers = (center_std * crng.standard_normal((n_centers, d))).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, n_centers, size=e - s)
        out[s:e] = centers[lab] + sigma * rng.standard_normal((e - s, d), dtype=np.float32)
    return out


def gamma_rows(n: int, d: int, seed: int, shape: float = 2.0, scale: float = 0.05,
               norm: float = 10.0, chunk: int = 1 << 16) -> np.ndarray:
    """Non-negative Gamma(shape, scale) rows, L2-normalised then scaled
    (the 1M x 960 NSG workload)."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = rng.gamma(shape, scale, size=(e - s, d)).astype(np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x * norm
    return out


def unit_rows(n: int, d: int, seed: int, chunk: int = 1 << 16) -> np.ndarray:
    """N(0,1) rows normalised to unit length (the 1M x 768 IP workload)."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = rng.standard_normal((e - s, d), dtype=np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x
    return out


def per_query_ef_k(nq: int, ef_lo: int = 16, ef_hi: int = 512, k_lo: int = 1, k_hi: int = 100,
                   seed: int = EFK_SEED):
    """ef log-uniform in [ef_lo, ef_hi] and k uniform in [k_lo, k_hi]."""
    rng = np.random.default_rng(seed)
    ef = np.rint(np.exp(rng.uniform(np.log(ef_lo), np.log(ef_hi), size=nq))).astype(np.int32)
    k = rng.integers(k_lo, k_hi + 1, size=nq).astype(np.int32)
    return ef, k
