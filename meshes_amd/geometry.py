"""Mesh geometries and seeded parameter draws used by the GPU path and its benchmarks.

Clements: L = N layers, shifts[l] = l % 2, lens[l] = (N - l % 2) // 2 (clements.py:56-57).
Butterfly / FFT: log2(N) layers with stride index l (gpu/test.py:42, gpu/fwdprop.cu:48-93).
Slot-layout parameters: p[L, N/2, n_p] with inactive slots ignored (gpu/gmem.cu:236-240).
"""
from __future__ import annotations

import numpy as np


def clements(N: int):
    l = np.arange(N)
    return (N - l % 2) // 2, l % 2


def butterfly_strides(N: int):
    n = int(np.log2(N))
    assert 1 << n == N, "butterfly meshes need N = 2^n"
    return np.arange(n)


def active_slots(N, L, lens, shifts):
    """Boolean [L, N/2] mask of active NN slots."""
    mask = np.zeros((L, N // 2), dtype=bool)
    for l in range(L):
        mask[l, shifts[l] // 2: shifts[l] // 2 + lens[l]] = True
    return mask


def uniform_phases(rng, L, N, n_p=2, dtype=np.float32):
    """Phases i.i.d. U[0, 2 pi) in slot layout."""
    return rng.uniform(0, 2 * np.pi, size=(L, N // 2, n_p)).astype(dtype)


def unit_cn(rng, B, N, real=False):
    """B i.i.d. CN(0, 1) vectors of length N, each normalised to unit norm, as a (B, N) array."""
    x = rng.standard_normal((B, N)) if real else rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    return x / np.linalg.norm(x, axis=1, keepdims=True)
