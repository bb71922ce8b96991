"""torch.autograd integration: the mesh as a differentiable function of its phases and inputs.

``mesh_apply(mesh, p, x)`` returns ``y = U(p) x`` for a batch ``x`` of shape (B, N) (one vector
per row) and differentiates through the hand-written adjoint kernel (backdiff) instead of
recording the layer loop.  The interface follows the reference's custom-VJP wrapper
(jax.py:15-36); torch's complex gradient convention (dL/dRe + i dL/dIm) is the meshes
convention, so no conjugation is needed at the boundary (jax.py:31-33 conjugates for JAX).

Phase screen (``is_phase``: the last N entries of p are phases phi applied as exp(i phi) at the
mesh input or output, mesh.py:38-41 and c/dot.cpp:97-143).  Forward: the screen kernel on a copy
of x before the mesh ("in") or on y after it ("out").  Backward: the screen's adjoint
(``meshprop_phase_screen`` with MESHPROP_OP_BACKDIFF) adds dJ/dphi_w = Im sum_b g conj(y) and maps
(field, gradient) back through the screen: before the mesh's adjoint for "out", after it for "in"
(where the forward saved the screened input S x).

Adjoint.  For unitary crossings (mzi, sym, orth) backdiff reconstructs the layer inputs by
inverse propagation, so the forward saves only its output.  For lossy crossings
(mzi_generic with gains) the forward stores the field entering every ``ckpt_every``-th
layer and the backward restarts each segment from the stored field (checkpointed adjoint),
with T^-1 for the reconstruction and T^H for the adjoint; with balanced arm gains
(T = c U) T^H = |det T| T^-1 and the unitary adjoint runs on the inverse rows alone.
"""
from __future__ import annotations

import torch

from . import _lib
from .mesh import MeshNetworkGPU


def _launch(mesh, op, B, p, dp, u_a, du_a, u_b, du_b, target=None, ckpt=None, ckpt_every=0):
    ldp, lds = mesh.N // 2 * mesh.X_np, mesh.N // 2 * mesh.X_ns
    g0, g1 = mesh._geom()
    s = mesh.p_splitter if mesh.p_splitter.numel() else None
    ws, nb = mesh._workspace(op, B)
    flags = mesh._adj_flags(mesh.p_splitter) if op == _lib.OP_BACKDIFF else 0
    _lib.run(op, mesh.X, mesh.N, mesh.L, B, mesh.fft, g0, g1, p, dp, ldp, s, lds, u_a, du_a, u_b, du_b, mesh.N,
             target=target, ckpt=ckpt, ckpt_every=ckpt_every, tile=mesh.tile, workspace=ws, ws_bytes=nb,
             flags=flags)


class MeshFunction(torch.autograd.Function):
    """y = U(p) x with x, y of shape (B, N); p flat (n_phase,) float32."""

    @staticmethod
    def forward(ctx, p, x, mesh: MeshNetworkGPU, ckpt_every: int):
        # The kernels read raw pointers: shapes, devices and lazy views are settled here.
        if x.ndim != 2 or x.shape[1] != mesh.N:
            raise ValueError(f"x must be (B, N={mesh.N}), got {tuple(x.shape)}")
        if p.numel() != mesh.n_phase:
            raise ValueError(f"p must hold n_phase={mesh.n_phase} values, got {p.numel()}")
        if x.device != mesh.device or p.device != mesh.device:
            raise ValueError(f"x ({x.device}) and p ({p.device}) must be on the mesh's device {mesh.device}")
        ctx.p_shape = p.shape
        p = p.detach().resolve_neg().to(torch.float32).reshape(-1).contiguous()
        xb = x.detach().resolve_conj().resolve_neg().to(mesh.dtype).contiguous()
        B = xb.shape[0]
        y = torch.empty_like(xb)
        ckpt = None
        if ckpt_every > 0 and mesh.L > 0:
            nck = (mesh.L + ckpt_every - 1) // ckpt_every
            ckpt = torch.empty((nck, B, mesh.N), dtype=xb.dtype, device=xb.device)
        n_m = mesh.n_cr * mesh.X_np  # mesh phases; the screen's N phases follow
        xs = torch.empty(0, device=xb.device)
        src = xb
        if mesh.is_phase and mesh.phi_pos == "in":  # exp(i phi) on a copy of the input
            xs = xb.clone()
            mesh._screen(_lib.OP_FWDPROP, B, p[n_m:], None, xs, None)
            src = xs
        _launch(mesh, _lib.OP_FWDPROP, B, p, None, src, None, y, None, ckpt=ckpt,
                ckpt_every=ckpt_every if ckpt is not None else 0)
        if mesh.is_phase and mesh.phi_pos == "out":
            mesh._screen(_lib.OP_FWDPROP, B, p[n_m:], None, y, None)
        ctx.mesh, ctx.ckpt_every = mesh, ckpt_every if ckpt is not None else 0
        ctx.save_for_backward(p, y, ckpt if ckpt is not None else torch.empty(0, device=xb.device), xs)
        return y

    @staticmethod
    def backward(ctx, gy):
        p, y, ckpt, xs = ctx.saved_tensors
        mesh = ctx.mesh
        B = y.shape[0]
        n_m = mesh.n_cr * mesh.X_np
        # A loss built from y.conj() hands in a lazy conjugate view: resolve it to plain data.
        g = gy.resolve_conj().resolve_neg().to(mesh.dtype).contiguous()
        dp = torch.zeros_like(p)
        if mesh.is_phase and mesh.phi_pos == "out":
            # dJ/dphi, then (y, g) back through the screen: the mesh's output and its gradient
            y = y.clone()  # the saved output is the caller's tensor
            if g.data_ptr() == gy.data_ptr():  # nor may the incoming gradient change
                g = g.clone()
            mesh._screen(_lib.OP_BACKDIFF, B, p[n_m:], dp[n_m:], y, g)
        gx = torch.empty_like(y)
        xr = torch.empty_like(y)
        _launch(mesh, _lib.OP_BACKDIFF, B, p, dp, y, g, xr, gx, ckpt=ckpt if ckpt.numel() else None,
                ckpt_every=ctx.ckpt_every)
        if mesh.is_phase and mesh.phi_pos == "in":
            # the screened input S x the forward saved, and the mesh's input gradient through the screen
            mesh._screen(_lib.OP_BACKDIFF, B, p[n_m:], dp[n_m:], xs.clone(), gx)
        return dp.view(ctx.p_shape), gx, None, None


def mesh_apply(mesh: MeshNetworkGPU, p: torch.Tensor, x: torch.Tensor, ckpt_every: int = 0) -> torch.Tensor:
    """Differentiable y = U(p) x (x: (B, N)), with the mesh's phase screen when it has one.
    ckpt_every > 0 stores forward checkpoints for a lossy mesh; for unitary crossings the adjoint
    reconstructs fields by inverse propagation."""
    return MeshFunction.apply(p, x, mesh, int(ckpt_every))


class MeshLayer(torch.nn.Module):
    """A mesh with trainable phases as a torch module: forward(x) = U(p) x for x of shape (B, N)."""

    def __init__(self, mesh: MeshNetworkGPU, ckpt_every: int = 0):
        super().__init__()
        self.mesh = mesh
        self.ckpt_every = ckpt_every
        self.p = torch.nn.Parameter(mesh.p_phase.clone())

    def forward(self, x):
        return mesh_apply(self.mesh, self.p, x, self.ckpt_every)
