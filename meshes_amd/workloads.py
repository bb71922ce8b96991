"""Synthetic workloads of the grouped and served forms (seeded; no datasets).

Heterogeneous mix: 64 meshes, half Clements (depth N) and half butterfly (log2 N stride
layers, gpu/test.py:42), N uniform over {32, 64, 128, 256, 512}, per-mesh batch log-uniform
in [1, 2048] complex64 unit-norm CN(0, 1) vectors, MZI phases U[0, 2 pi), one fwd+grad job per
mesh with the loss J = sum |U x - t|^2 toward a random unit-norm target t.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np
import torch

from . import geometry
from .batched import GroupJob
from .mesh import MeshNetworkGPU

MIX_N = (32, 64, 128, 256, 512)


def mesh_of(rng, topo: str, N: int, X: str = "mzi") -> MeshNetworkGPU:
    n_p = 2
    if topo == "clements":
        lens, shifts = geometry.clements(N)
        p = geometry.uniform_phases(rng, N, N, n_p)
        return MeshNetworkGPU(N, N, lens, shifts, p, np.zeros(0), X, is_phase=False)
    st = geometry.butterfly_strides(N)
    L = len(st)
    p = geometry.uniform_phases(rng, L, N, n_p)
    return MeshNetworkGPU(N, L, np.full(L, N // 2), np.zeros(L, int), p, np.zeros(0), X, is_phase=False,
                          strides=st)


def mix_specs(seed: int = 0, n_mesh: int = 64, bmax: int = 2048) -> List[Tuple[str, int, int]]:
    rng = np.random.RandomState(seed)
    topo = ["clements"] * (n_mesh // 2) + ["butterfly"] * (n_mesh - n_mesh // 2)
    rng.shuffle(topo)
    out = []
    for t in topo:
        N = int(rng.choice(MIX_N))
        B = int(np.exp(rng.uniform(0.0, np.log(bmax + 1.0))))
        out.append((t, N, min(max(B, 1), bmax)))
    return out


def mix_jobs(specs, seed: int = 1, device=None) -> List[GroupJob]:
    rng = np.random.RandomState(seed)
    dev = device or torch.device("cuda", torch.cuda.current_device())
    jobs = []
    for topo, N, B in specs:
        m = mesh_of(rng, topo, N)
        x = torch.as_tensor(geometry.unit_cn(rng, B, N).astype(np.complex64), device=dev)
        t = torch.as_tensor(geometry.unit_cn(rng, B, N).astype(np.complex64), device=dev)
        jobs.append(GroupJob(m, "grad", x=x, target=t))
    return jobs


def updates(jobs) -> int:
    """MZI-vector updates of one fwd+grad pass: active crossings x vectors."""
    tot = 0
    for j in jobs:
        m = j.mesh
        if m.fft:
            tot += j.B * m.L * (m.N // 2)
        else:
            tot += j.B * int(m.lens.sum().item())
    return tot


def time_ms(fn, steps: int, warmup: int) -> List[float]:
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        out.append(a.elapsed_time(b))
    return out
