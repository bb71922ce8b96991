"""Meshes with arbitrary waveguide permutations between layers, on the GPU.

A StructuredMeshNetwork may carry a permutation (gather, v <- v[p]) before every layer and one
after the last (``perm``, L + 1 entries, mesh.py:240-290); the CPU engine applies them between
layers (c/dot.cpp:150-152, 289-291, mesh.py:47-48, 71-72) and the reference GPU path rejects
them (mesh.py:573).  Permutations that only pair waveguides at a power-of-two stride map onto
the FFT geometry inside the kernels (``bridge.strides_from_perms``); any other one splits the
mesh here into segments of consecutive layers with no permutation between them.  Each segment
is a ``MeshNetworkGPU`` (NN geometry, its slice of the slot-layout parameters), and the
permutations run between segments as the ``meshprop_permute`` gather kernel.  The adjoint walks
the segments backwards, undoing each permutation with its inverse on the field and on the
gradient (as mesh.py:381 passes the inverse permutations to the CPU engine's backward mode).
The phase screen, when present, sits before the first permutation (``phi_pos='in'``) or after
the last (``'out'``), as in the CPU engine.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .mesh import MeshNetworkGPU, _device, as_field, as_params, numel


def segment_bounds(L: int, perms: Sequence[Optional[np.ndarray]]) -> List[int]:
    """Layer indices that start a segment, plus L: 0, every 0 < n < L with a permutation before
    layer n, and L."""
    return [0] + [n for n in range(1, L) if perms[n] is not None] + [L]


class PermutedMeshGPU:
    """A mesh of NN crossing layers with arbitrary permutations between them (the
    StructuredMeshNetwork form of mesh.py:240-290), with the MeshNetworkGPU call interface:
    ``dot`` (forward and forward-mode), ``grad_phi`` (reverse mode), ``matrix``, ``L2norm``,
    ``grad_phi_target``, ``hvp``/``hvp_target``/``hess``/``hess_target`` (central differences of
    grad_phi, as gpu/mesh.py:236-288).

    :param perms: L + 1 entries, None or a permutation of range(N); ``perms[n]`` is applied before
        layer n, ``perms[L]`` after the last.
    Other parameters as MeshNetworkGPU (slot-layout p_phase of size L*(N/2)*X_np, plus N screen
    phases when is_phase)."""

    def __init__(self, N: int, L: int, lens, shifts, perms, p_phase, p_splitter, X: str, phi_pos="out",
                 is_phase=True, tile=None):
        _lib.load()
        assert len(perms) == L + 1, "one permutation slot before every layer and one after the last"
        assert phi_pos in ["in", "out"] and not (is_phase and X == "orth")
        self._N, self._L = int(N), int(L)
        self.X, self.phi_pos, self.is_phase, self.tile = X, phi_pos, bool(is_phase), tile
        self.X_np, self.X_ns = _lib.N_PHASE[X], _lib.N_SPLIT[X]
        self.device = _device()
        self.dtype = torch.float32 if X == "orth" else torch.complex64
        assert numel(p_phase) == self.n_phase and numel(p_splitter) in [self.n_splitter, 0]
        self.p_phase = as_params(p_phase, self.n_phase, self.device)
        self.p_splitter = (as_params(p_splitter, self.n_splitter, self.device) if numel(p_splitter)
                           else torch.empty(0, dtype=torch.float32, device=self.device))
        lens = np.asarray(lens, np.int64).reshape(L)
        shifts = np.asarray(shifts, np.int64).reshape(L)
        self.perm, self._fwd, self._inv = [], [], []
        for q in perms:
            if q is None:
                self.perm.append(None)
                self._fwd.append(None)
                self._inv.append(None)
                continue
            q = np.asarray(q, np.int64).reshape(-1)
            assert q.size == N and np.array_equal(np.sort(q), np.arange(N)), "each entry must permute range(N)"
            self.perm.append(q)
            self._fwd.append(torch.as_tensor(q.astype(np.int32), device=self.device))
            self._inv.append(torch.as_tensor(np.argsort(q).astype(np.int32), device=self.device))
        self.bounds = segment_bounds(L, self.perm)
        ldp, lds = self._ld()
        p0 = self.p_phase[:self.n_cr * self.X_np]
        self.segments = []
        for a, b in zip(self.bounds[:-1], self.bounds[1:]):
            s0 = self.p_splitter[a * lds:b * lds] if self.p_splitter.numel() else np.zeros(0)
            self.segments.append((a, b, MeshNetworkGPU(N, b - a, lens[a:b], shifts[a:b], p0[a * ldp:b * ldp], s0, X,
                                                       is_phase=False, tile=tile)))

    # ---- properties (as MeshNetworkGPU) -----------------------------------------------
    @property
    def N(self):
        return self._N

    @property
    def M(self):
        return self._N

    @property
    def L(self):
        return self._L

    @property
    def n_cr(self):
        return self.L * (self.N // 2)

    @property
    def n_phase(self):
        return self.n_cr * self.X_np + self.is_phase * self.N

    @property
    def n_splitter(self):
        return self.n_cr * self.X_ns

    def _ld(self):
        return self.N // 2 * self.X_np, self.N // 2 * self.X_ns

    _defaults = MeshNetworkGPU._defaults
    _L2norm_fn = MeshNetworkGPU._L2norm_fn
    L2norm = MeshNetworkGPU.L2norm
    grad_phi_target = MeshNetworkGPU.grad_phi_target
    _hvp_difference = MeshNetworkGPU._hvp_difference
    hess = MeshNetworkGPU.hess
    hess_target = MeshNetworkGPU.hess_target

    def exact_hvp_supported(self) -> bool:
        return False

    def hvp(self, vec, v, w, p_phase=None, p_splitter=None, eps=1e-3):
        """H vec by central differences of grad_phi (gpu/mesh.py:236-251)."""
        p = self.p_phase if p_phase is None else as_params(p_phase, self.n_phase, self.device)
        return self._hvp_difference(as_params(vec, self.n_phase, self.device), v, w, p, p_splitter, eps)

    def hvp_target(self, vec, U_target, p_phase=None, p_splitter=None, eps=1e-3):
        J = [0]
        w = self._L2norm_fn(J)
        Ut = torch.as_tensor(np.asarray(U_target) if not isinstance(U_target, torch.Tensor) else U_target)
        return self.hvp(vec, Ut.to(self.dtype).to(self.device).T.conj(), w, p_phase, p_splitter, eps)

    def matrix(self, p_phase=None, p_splitter=None) -> torch.Tensor:
        return self.dot(torch.eye(self.N, dtype=self.dtype, device=self.device), p_phase, p_splitter)

    # ---- pieces -----------------------------------------------------------------------
    def _gather(self, u, idx):
        out = torch.empty_like(u)
        _lib.permute(self.N, u.shape[0], idx, u, out, self.N)
        return out

    def _screen(self, op, phi, dphi, u, du):
        B = u.shape[0]
        ws, nb = None, 0
        if op == _lib.OP_BACKDIFF:
            nb = _lib.phase_screen_workspace_size(self.N, B)
            ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.device)
        _lib.phase_screen(op, self.N, B, phi, dphi, u, du, self.N, workspace=ws, ws_bytes=nb)

    def _slices(self, p, s, a, b):
        ldp, lds = self._ld()
        return p[a * ldp:b * ldp], (s[a * lds:b * lds] if s.numel() else torch.empty(0, device=self.device))

    # ---- propagation ------------------------------------------------------------------
    def dot(self, v, p_phase=None, p_splitter=None, dp=None, dv=None):
        """out = U v, or (out, dout) with dout = U dv + (dU/dp . dp) v; (N, B) in and out."""
        p, s = self._defaults(p_phase, p_splitter)
        u = as_field(v, self.N, self.dtype, self.device).clone()
        jvp = dp is not None or dv is not None
        n_cr_p = self.n_cr * self.X_np
        if jvp:
            dpt = as_params(dp, self.n_phase, self.device) if dp is not None else torch.zeros_like(p)
            du = (as_field(dv, self.N, self.dtype, self.device).clone() if dv is not None else torch.zeros_like(u))
        if self.is_phase and self.phi_pos == "in":
            if jvp:
                self._screen(_lib.OP_FWDDIFF, p[n_cr_p:], dpt[n_cr_p:], u, du)
            else:
                self._screen(_lib.OP_FWDPROP, p[n_cr_p:], None, u, None)
        for a, b, seg in self.segments:
            if self._fwd[a] is not None:
                u = self._gather(u, self._fwd[a])
                if jvp:
                    du = self._gather(du, self._fwd[a])
            ps, ss = self._slices(p, s, a, b)
            if jvp:
                y, dy = seg.dot(u.T, ps, ss, dp=self._slices(dpt, s, a, b)[0], dv=du.T)
                u, du = y.T, dy.T
            else:
                u = seg.dot(u.T, ps, ss).T
        if self._fwd[self.L] is not None:
            u = self._gather(u, self._fwd[self.L])
            if jvp:
                du = self._gather(du, self._fwd[self.L])
        if self.is_phase and self.phi_pos == "out":
            if jvp:
                self._screen(_lib.OP_FWDDIFF, p[n_cr_p:], dpt[n_cr_p:], u, du)
            else:
                self._screen(_lib.OP_FWDPROP, p[n_cr_p:], None, u, None)
        return (u.T, du.T) if jvp else u.T

    def grad_phi(self, v, dJdv, p_phase=None, p_splitter=None, out_dp=None, out_dJdv=None, vpos="in"):
        """Backpropagation: (out_dp, dJ/dv_in), out_dp accumulated into when given; v the input
        (vpos='in', propagated forward first) or the output (vpos='out'); dJdv an (N, B) array or
        a callable of the output."""
        assert vpos in ["in", "out"]
        p, s = self._defaults(p_phase, p_splitter)
        vout = self.dot(v, p, s) if vpos == "in" else v
        u = as_field(vout, self.N, self.dtype, self.device).clone()
        if callable(dJdv):
            dJdv = dJdv(u.T)
        g = as_field(dJdv, self.N, self.dtype, self.device).clone()
        if out_dp is None:
            out_dp = torch.zeros(self.n_phase, dtype=torch.float32, device=self.device)
        assert isinstance(out_dp, torch.Tensor) and out_dp.dtype == torch.float32 and out_dp.is_contiguous()
        assert out_dp.numel() == self.n_phase and out_dp.device == self.device
        n_cr_p = self.n_cr * self.X_np
        ldp, _ = self._ld()
        if self.is_phase and self.phi_pos == "out":
            self._screen(_lib.OP_BACKDIFF, p[n_cr_p:], out_dp[n_cr_p:], u, g)
        if self._inv[self.L] is not None:
            u, g = self._gather(u, self._inv[self.L]), self._gather(g, self._inv[self.L])
        for a, b, seg in reversed(self.segments):
            ps, ss = self._slices(p, s, a, b)
            gin = torch.empty_like(g)
            u = seg._backdiff(u, g, ps, ss, out_dp[a * ldp:b * ldp], gin)
            g = gin
            if self._inv[a] is not None:
                u, g = self._gather(u, self._inv[a]), self._gather(g, self._inv[a])
        if self.is_phase and self.phi_pos == "in":
            self._screen(_lib.OP_BACKDIFF, p[n_cr_p:], out_dp[n_cr_p:], u, g)
        if out_dJdv is None:
            return out_dp, g.T
        out_dJdv.copy_(g.T)
        return out_dp, out_dJdv
