"""From the reference's CPU mesh objects to the GPU path, and the optimiser that drives it.

``to_gpu(net)`` replaces ``StructuredMeshNetwork.gpu()`` (mesh.py:568-579): it reads a
StructuredMeshNetwork-like object through its attributes only (``N``, ``L``, ``lens``,
``shifts``, ``p_phase``, ``p_splitter``, ``X`` with ``n_phase``/``n_splitter``, ``perm``,
``phi_pos``, ``is_phase``) and builds a ``MeshNetworkGPU``.  Beyond the reference's bridge it

* keeps the phase screen (the reference passes ``is_phase`` but drops ``phi_out``, so its
  bridge only works with ``is_phase=False``; here the screen runs on the GPU);
* maps meshes with permutations whose layers are stride-2^s butterflies -- e.g.
  ``ButterflyNetwork`` (butterfly.py:46-51), N-1 full layers with gathers around the
  stride-s layers -- onto the GPU's FFT geometry (strides per layer, slot order), where the
  reference asserts that there are no permutations (mesh.py:573).

``grid_phase`` / ``grid_splitter`` restate mesh.py:513-536 (packed crossing parameters onto the
(L, N/2, n) slot grid; inactive slots NaN).  ``MeshOptimizer`` restates gpu/opt.py:16-69 on torch
tensors: scipy.optimize.minimize over the crossing phases of ``|offdiag(U_mesh U^H)|^2``.
"""
from __future__ import annotations

from time import time
from typing import List, Optional, Tuple

import numpy as np
import torch

from .mesh import MeshNetworkGPU

# Crossing classes of crossing.py with a GPU model (mesh.py:576 maps the first two).
X_NAMES = {"MZICrossing": "mzi", "SymCrossing": "sym", "MZICrossingGeneric": "mzi_generic"}


def _packed(net) -> Tuple[np.ndarray, np.ndarray, Optional[np.ndarray]]:
    """(p_crossing (n_cr, n_p), p_splitter (n_cr, n_s), phi_out or None) of a StructuredMeshNetwork
    (mesh.py:307-332, 335-341)."""
    n_p, n_s = int(net.X.n_phase), int(net.X.n_splitter)
    n_cr = int(sum(net.lens))
    p = np.asarray(net.p_phase, dtype=np.float64).reshape(-1)
    p_cr = p[:n_cr * n_p].reshape(n_cr, n_p)
    phi = p[n_cr * n_p:] if net.is_phase else None
    s = np.asarray(net.p_splitter, dtype=np.float64) * np.ones((n_cr, n_s))
    return p_cr, s, phi


def _grid(values: np.ndarray, N: int, lens, shifts, n: int) -> np.ndarray:
    L = len(lens)
    out = np.full((L, (N + 1) // 2, n), np.nan)
    ind = 0
    for l, (j0, nj) in enumerate(zip(shifts, lens)):
        out[l, j0 // 2:j0 // 2 + nj, :] = values[ind:ind + nj]
        ind += nj
    return out


def grid_phase(net) -> np.ndarray:
    """Crossing parameters on the (L, ceil(N/2), n_phase) slot grid, NaN where no crossing
    (mesh.py:513-522)."""
    p_cr, _, _ = _packed(net)
    return _grid(p_cr, net.N, net.lens, net.shifts, int(net.X.n_phase))


def grid_splitter(net) -> np.ndarray:
    """Splitter parameters on the slot grid (mesh.py:524-536)."""
    if not np.iterable(net.p_splitter):
        return np.zeros([net.L, 0])
    _, s, _ = _packed(net)
    return _grid(s, net.N, net.lens, net.shifts, int(net.X.n_splitter))


def strides_from_perms(N: int, lens, shifts, perms) -> Tuple[np.ndarray, List[np.ndarray]]:
    """Stride index and slot order of every layer of a permuted mesh.

    The gathers are composed into the map from positions to waveguides (``perms[n]`` before
    layer n, ``perms[L]`` after the last, as c/dot.cpp:150-152, 289-291 apply them).  Crossing k
    of layer n acts on the waveguides at positions (shift + 2k, shift + 2k + 1); the layer maps
    onto the FFT geometry when every such pair is (i, i + S) with i & S == 0 for one S = 2^s and
    all N/2 pairs are present.  Returns (strides[L], slot[L][lens[n]]): crossing k of layer n is
    FFT slot slot[n][k] = (i // 2S) * S + i % 2S.  Raises ValueError otherwise, or when the last
    gather does not restore the waveguide order."""
    L = len(lens)
    pos = np.arange(N)  # waveguide held at each position
    strides, slots = [], []
    for n in range(L):
        if perms[n] is not None:
            pos = pos[np.asarray(perms[n])]
        if lens[n] != N // 2 or shifts[n] != 0:
            raise ValueError(f"layer {n}: a permuted mesh maps onto the GPU only with full layers")
        a = pos[0:N:2]
        b = pos[1:N:2]
        S = b - a
        if not (np.all(S == S[0]) and S[0] > 0 and (S[0] & (S[0] - 1)) == 0 and np.all((a & S[0]) == 0)):
            raise ValueError(f"layer {n} is not a stride-2^s butterfly layer")
        S = int(S[0])
        strides.append(S.bit_length() - 1)
        slots.append((a // (2 * S)) * S + a % (2 * S))
    if perms[L] is not None:
        pos = pos[np.asarray(perms[L])]
    if not np.array_equal(pos, np.arange(N)):
        raise ValueError("the permutations do not restore the waveguide order at the output")
    return np.asarray(strides, dtype=np.int64), slots


def to_gpu(net, tile=None):
    """MeshNetworkGPU of a StructuredMeshNetwork-like object (the replacement of mesh.py:568-579).

    Permutations that pair waveguides at power-of-two strides map onto the FFT geometry; any
    other permutations give a PermutedMeshGPU (segments of layers between gathers, with the same
    call interface), where the reference GPU path rejects them (mesh.py:573)."""
    name = type(net.X).__name__
    if name not in X_NAMES:
        raise NotImplementedError(f"no GPU model for crossing {name}")
    X = X_NAMES[name]
    N, L = int(net.N), int(net.L)
    if N % 2:
        raise ValueError("the GPU path needs an even N")
    lens, shifts = [int(x) for x in net.lens], [int(x) for x in net.shifts]
    n_p, n_s = int(net.X.n_phase), int(net.X.n_splitter)
    p_cr, s_cr, phi = _packed(net)
    perm = list(net.perm) if getattr(net, "perm", None) is not None else [None] * (L + 1)
    splitter = np.iterable(net.p_splitter) or np.any(np.asarray(net.p_splitter) != 0)
    butterfly = True
    if any(q is not None for q in perm):
        try:
            strides, slots = strides_from_perms(N, lens, shifts, perm)
        except ValueError:
            butterfly = False  # any other permutations: segments between gathers
    if all(q is None for q in perm) or not butterfly:
        p_slot = np.nan_to_num(_grid(p_cr, N, lens, shifts, n_p))
        s_slot = np.nan_to_num(_grid(s_cr, N, lens, shifts, n_s)) if splitter else np.zeros(0)
        kw = {}
        if not butterfly:
            from .permuted import PermutedMeshGPU
            p_full = p_slot.astype(np.float32).reshape(-1)
            if net.is_phase:
                p_full = np.concatenate([p_full, np.asarray(phi, np.float32)])
            return PermutedMeshGPU(N, L, np.asarray(lens), np.asarray(shifts), perm, p_full,
                                   np.asarray(s_slot, np.float32).reshape(-1), X, phi_pos=net.phi_pos,
                                   is_phase=bool(net.is_phase), tile=tile)
    else:
        p_slot = np.zeros((L, N // 2, n_p))
        s_slot = np.zeros((L, N // 2, n_s))
        ind = 0
        for n in range(L):
            p_slot[n, slots[n]] = p_cr[ind:ind + lens[n]]
            s_slot[n, slots[n]] = s_cr[ind:ind + lens[n]]
            ind += lens[n]
        if not splitter:
            s_slot = np.zeros(0)
        kw = dict(strides=strides)
    p_full = p_slot.astype(np.float32).reshape(-1)
    if net.is_phase:
        p_full = np.concatenate([p_full, np.asarray(phi, np.float32)])
    return MeshNetworkGPU(N, L, np.asarray(lens), np.asarray(shifts), p_full, np.asarray(s_slot, np.float32).reshape(-1),
                          X, phi_pos=net.phi_pos, is_phase=bool(net.is_phase), tile=tile, **kw)


class MeshOptimizer:
    """scipy.optimize.minimize over the crossing phases of a mesh on the GPU (gpu/opt.py:16-69).

    The objective is ``|offdiag(U_mesh U^H)|^2`` through ``MeshNetworkGPU.grad_phi_target``
    (the phase screen is left out, as gpu/opt.py:27 does); ``p`` holds the packed crossing
    phases of the CPU mesh, mapped onto the GPU slots."""

    def __init__(self, mesh, U):
        is_phase = mesh.is_phase
        try:
            mesh.is_phase = False
            self.mesh = to_gpu(mesh)
        finally:
            mesh.is_phase = is_phase
        n_p = int(mesh.X.n_phase)
        p_cr, _, _ = _packed(mesh)
        self.p = np.copy(p_cr.reshape(-1))
        self.U = torch.as_tensor(np.asarray(U), dtype=torch.complex64, device=self.mesh.device)
        self._Jcache = [None, 0.0]
        # packed crossing k -> flat slot index of the GPU mesh
        N, lens, shifts = int(mesh.N), list(mesh.lens), list(mesh.shifts)
        perm = list(mesh.perm) if getattr(mesh, "perm", None) is not None else [None] * (len(lens) + 1)
        from .permuted import PermutedMeshGPU
        if all(q is None for q in perm) or isinstance(self.mesh, PermutedMeshGPU):
            # NN slots (also for a PermutedMeshGPU: its segments keep the NN geometry)
            self._inds = np.concatenate([np.arange(n * (N // 2) + s // 2, n * (N // 2) + s // 2 + k)
                                         for n, (s, k) in enumerate(zip(shifts, lens))])
        else:
            _, slots = strides_from_perms(N, lens, shifts, perm)
            self._inds = np.concatenate([n * (N // 2) + np.asarray(sl) for n, sl in enumerate(slots)])
        self._inds_t = torch.as_tensor(self._inds, device=self.mesh.device)
        self._n_p = n_p
        self.yt: list = []
        self.t: list = []

    def _to_gpu(self, p):
        out = torch.zeros((self.mesh.n_cr, self._n_p), dtype=torch.float32, device=self.mesh.device)
        out[self._inds_t] = torch.as_tensor(np.asarray(p, np.float32).reshape(-1, self._n_p), device=self.mesh.device)
        return out.reshape(-1)

    def _from_gpu(self, dp):
        return dp.reshape(self.mesh.n_cr, self._n_p)[self._inds_t].reshape(-1).cpu().numpy()

    def f(self, p):
        key = np.asarray(p).tobytes()
        if self._Jcache[0] == key:
            return self._Jcache[1]
        out = self.mesh.dot(self.U.T.conj(), self._to_gpu(p))
        out = out - torch.diag_embed(torch.diagonal(out))
        return float(torch.linalg.norm(out) ** 2)

    def jac(self, p):
        J, dp = self.mesh.grad_phi_target(self.U, self._to_gpu(p))
        self._Jcache = [np.asarray(p).tobytes(), float(J)]
        return self._from_gpu(dp)

    def run(self, method, p=None, tol=1e-5, maxiter=10000, tick=-1):
        from scipy.optimize import minimize
        p = self.p if p is None else p
        self.yt, self.t = [], []
        t0 = time()

        def f(x):
            out = self.f(x)
            self.yt.append(out)
            self.t.append(time() - t0)
            if tick > 0 and len(self.yt) % tick == 0:
                print(".", end="")
            return np.float64(out)

        def jac(x):
            return self.jac(x).astype(np.float64)

        out = minimize(fun=f, x0=np.asarray(p, np.float64), jac=jac, method=method, tol=tol,
                       options={"maxiter": maxiter})
        if tick > 0:
            print()
        self.yt = out.yt = np.sqrt(np.asarray(self.yt)) / np.sqrt(len(self.U))
        self.t = out.t = np.asarray(self.t)
        return out
