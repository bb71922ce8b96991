"""MeshNetworkGPU: the drop-in for gpu/mesh.py of the reference, on MI355X.

Same constructor, properties and methods as the reference class (gpu/mesh.py:21-297);
arrays are torch tensors on the current HIP device instead of CuPy arrays, and every
propagation runs the hand-written gfx950 kernels of libmeshprop.so through its C ABI.

Field arrays follow the reference convention: a batch of B vectors of N modes is an
``(N, B)`` array in Fortran order (vector-contiguous).  In torch that is the transpose
view of a contiguous ``(B, N)`` tensor; returned fields are such views.  Inputs may be
torch tensors (any device, layout) or NumPy arrays; they are copied only when their
dtype, device or layout differ.
"""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
import torch

from . import _lib

_DTYPE = {"mzi": torch.complex64, "sym": torch.complex64, "orth": torch.float32, "mzi_generic": torch.complex64}


def _device():
    if not torch.cuda.is_available():
        raise _lib.MeshpropUnavailable("no HIP device is visible; the mesh kernels run on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


def as_field(v, N, dtype, device) -> torch.Tensor:
    """(N, B) field -> contiguous (B, N) tensor (the Fortran buffer), copying only if needed."""
    if not isinstance(v, torch.Tensor):
        v = torch.as_tensor(np.asarray(v))
    if v.ndim == 1:
        v = v.reshape(N, 1)
    assert v.ndim == 2 and v.shape[0] == N, f"field must be (N={N}, B), got {tuple(v.shape)}"
    t = v.T.resolve_conj().resolve_neg()  # lazy conjugate/negative views hold the plain data
    if t.dtype != dtype:
        t = t.to(dtype)
    if t.device != device:
        t = t.to(device)
    return t.contiguous()


def numel(x) -> int:
    return x.numel() if isinstance(x, torch.Tensor) else int(np.size(x))


def as_params(x, n, device) -> torch.Tensor:
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(np.asarray(x, dtype=np.float32))
    x = x.to(device=device, dtype=torch.float32).reshape(-1).contiguous()
    assert x.numel() == n, f"expected {n} parameters, got {x.numel()}"
    return x


class MeshNetworkGPU:
    """GPU mesh of 2x2 crossings (gpu/mesh.py:21-93).

    :param N: number of modes (even, N <= 1024).
    :param L: number of layers.
    :param lens, shifts: NN geometry, crossings per layer and their offset (ignored if ``strides`` is given).
    :param p_phase: crossing parameters in slot layout, size L*(N/2)*X_np.
    :param p_splitter: splitter imperfections, size L*(N/2)*X_ns or 0.
    :param X: crossing type: 'mzi', 'sym', 'orth' (the reference's three) or 'mzi_generic' (lossy MZI).
    :param phi_pos, is_phase: the phase screen exp(i phi) of the last N entries of p_phase, at the
        output ('out') or input ('in') of the mesh (mesh.py:38-41, c/dot.cpp:97-143).  The reference
        GPU path rejects is_phase (gpu/mesh.py:65); here it runs as a separate HIP kernel
        (meshprop_phase_screen).
    :param strides: optional FFT/butterfly geometry (stride index per layer); the reference only
        reaches these kernels from gpu/test.py.
    :param tile: optional (w, v, waves, chunk) tile override passed to the library.
    """

    def __init__(self, N: int, L: int, lens, shifts, p_phase, p_splitter, X: str, phi_pos="out", is_phase=True,
                 strides=None, tile=None):
        _lib.load()
        assert phi_pos in ["in", "out"]
        assert not (is_phase and X == "orth"), "the phase screen needs complex fields"
        assert X in ["mzi", "sym", "orth", "mzi_generic"]
        assert N % 2 == 0
        assert N <= 1024, "meshes wider than 1024 modes are not supported"
        self._N, self._L = int(N), int(L)
        self.X = X
        self.phi_pos = phi_pos
        self.is_phase = is_phase
        self.dtype = _DTYPE[X]
        self.X_ns = _lib.N_SPLIT[X]
        self.X_np = _lib.N_PHASE[X]
        self.device = _device()
        self.tile = tile
        assert numel(p_phase) == self.n_phase
        assert numel(p_splitter) in [self.n_splitter, 0]
        self.p_phase = as_params(p_phase, self.n_phase, self.device)
        self.p_splitter = (as_params(p_splitter, self.n_splitter, self.device) if numel(p_splitter)
                           else torch.empty(0, dtype=torch.float32, device=self.device))
        self.fft = strides is not None
        if self.fft:
            st = np.asarray(strides, dtype=np.int64).reshape(L)
            assert np.all((st >= 0) & (N % (2 << st) == 0)), "each stride 2^s must tile N"
            self.strides = torch.as_tensor(st.astype(np.int32), device=self.device)
            self.lens = torch.full((L,), N // 2, dtype=torch.int32, device=self.device)
            self.shifts = torch.zeros(L, dtype=torch.int32, device=self.device)
        else:
            ln = np.asarray(lens, dtype=np.int64).reshape(L)
            sh = np.asarray(shifts, dtype=np.int64).reshape(L)
            assert np.all(ln >= 0) and np.all(sh >= 0) and np.all(sh + 2 * ln <= N), "crossing beyond the mesh"
            self.lens = torch.as_tensor(ln.astype(np.int32), device=self.device)
            self.shifts = torch.as_tensor(sh.astype(np.int32), device=self.device)
            self.strides = None
        self._ws = {}
        self._balanced = {}  # balanced-gain check of the splitter tensor (generic MZI)

    # ---- properties (gpu/mesh.py:95-137) ----------------------------------------------
    @property
    def L(self):
        return self._L

    @property
    def M(self):
        return self._N

    @property
    def N(self):
        return self._N

    @property
    def n_splitter(self):
        return self.n_cr * self.X_ns

    @property
    def n_phase(self):
        return self.n_cr * self.X_np + self.is_phase * self.N

    @property
    def n_cr(self):
        """Number of crossing slots, including blank (inactive) ones."""
        return self.L * (self.N // 2)

    @property
    def p_crossing(self):
        """Crossing parameters as an (L, N/2, X_np) view of p_phase."""
        return self.p_phase[:self.n_cr * self.X_np].reshape(self.L, self.N // 2, self.X_np)

    @p_crossing.setter
    def p_crossing(self, p):
        self.p_crossing[:] = torch.as_tensor(p, dtype=torch.float32, device=self.device).reshape(
            self.L, self.N // 2, self.X_np)

    @property
    def phi_out(self):
        return self.p_phase[self.n_cr * self.X_np:]

    # ---- helpers ----------------------------------------------------------------------
    def _geom(self):
        return (self.strides, None) if self.fft else (self.lens, self.shifts)

    def _defaults(self, p_phase, p_splitter):
        p = self.p_phase if p_phase is None else as_params(p_phase, self.n_phase, self.device)
        if p_splitter is None:
            s = self.p_splitter
        else:
            s = (as_params(p_splitter, self.n_splitter, self.device) if numel(p_splitter)
                 else torch.empty(0, dtype=torch.float32, device=self.device))
        assert p.numel() == self.n_phase and s.numel() in [self.n_splitter, 0]
        return p, s

    def _workspace(self, op, B):
        key = (op, B)
        nbytes = _lib.workspace_size(op, self.X, self.N, self.L, B, self.fft, self.tile)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
            self._ws[key] = ws
        return ws, nbytes

    def _ld(self):
        return self.N // 2 * self.X_np, self.N // 2 * self.X_ns

    def _adj_flags(self, s) -> int:
        """MESHPROP_SCALED_UNITARY for a generic MZI whose arm gains are balanced in every
        crossing (g11 = g21, g12 = g22: T = c U with U unitary, since the generic splitters and
        phase errors are lossless), 0 otherwise.  Such a mesh runs the unitary adjoint on one
        inverse-row table with T^H = |det T| T^-1.  Checked on the device once per splitter
        tensor version (one synchronisation)."""
        if self.X != "mzi_generic":
            return 0
        if s.numel() == 0:
            return _lib.SCALED_UNITARY
        key = (s.data_ptr(), s._version, s.numel())
        hit = self._balanced.get(key)
        if hit is None:
            q = s.view(-1, 14)
            hit = bool(torch.logical_and(q[:, 10] == q[:, 11], q[:, 12] == q[:, 13]).all().item())
            self._balanced = {key: hit}
        return _lib.SCALED_UNITARY if hit else 0

    # ---- propagation (gpu/mesh.py:144-174) --------------------------------------------
    def matrix(self, p_phase=None, p_splitter=None) -> torch.Tensor:
        return self.dot(torch.eye(self.N, dtype=self.dtype, device=self.device), p_phase, p_splitter)

    def _screen(self, op, B, phi, dphi, u, du):
        ws = None
        nb = 0
        if op == _lib.OP_BACKDIFF:
            nb = _lib.phase_screen_workspace_size(self.N, B)
            ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.device)
        _lib.phase_screen(op, self.N, B, phi, dphi, u, du, self.N, workspace=ws, ws_bytes=nb)

    def dot(self, v, p_phase=None, p_splitter=None, dp=None, dv=None, out=None, dout=None):
        """out = U v (fwdprop), or (out, dout) with dout = U dv + (dU/dp . dp) v (fwddiff)."""
        vb = as_field(v, self.N, self.dtype, self.device)
        B = vb.shape[0]
        p, s = self._defaults(p_phase, p_splitter)
        ldp, lds = self._ld()
        g0, g1 = self._geom()
        ob = self._out_buffer(out, B)
        n_cr_p = self.n_cr * self.X_np
        screen_in = self.is_phase and self.phi_pos == "in"
        if dp is None and dv is None:
            if screen_in:  # exp(i phi) on a copy of the input, then the mesh in place
                ob.copy_(vb)
                self._screen(_lib.OP_FWDPROP, B, p[n_cr_p:], None, ob, None)
                vb = ob
            ws, nb = self._workspace(_lib.OP_FWDPROP, B)
            _lib.run(_lib.OP_FWDPROP, self.X, self.N, self.L, B, self.fft, g0, g1, p, None, ldp,
                     s if s.numel() else None, lds, vb, None, ob, None, self.N, tile=self.tile, workspace=ws,
                     ws_bytes=nb)
            if self.is_phase and not screen_in:
                self._screen(_lib.OP_FWDPROP, B, p[n_cr_p:], None, ob, None)
            return self._result(ob, out)
        dob = self._out_buffer(dout, B)
        dpt = as_params(dp, self.n_phase, self.device) if dp is not None else torch.zeros_like(p)
        dvb = as_field(dv, self.N, self.dtype, self.device) if dv is not None else None
        if screen_in:
            ob.copy_(vb)
            if dvb is None:
                dob.zero_()
            else:
                dob.copy_(dvb)
            self._screen(_lib.OP_FWDDIFF, B, p[n_cr_p:], dpt[n_cr_p:], ob, dob)
            vb, dvb = ob, dob
        ws, nb = self._workspace(_lib.OP_FWDDIFF, B)
        _lib.run(_lib.OP_FWDDIFF, self.X, self.N, self.L, B, self.fft, g0, g1, p, dpt, ldp,
                 s if s.numel() else None, lds, vb, dvb, ob, dob, self.N, tile=self.tile, workspace=ws, ws_bytes=nb)
        if self.is_phase and not screen_in:
            self._screen(_lib.OP_FWDDIFF, B, p[n_cr_p:], dpt[n_cr_p:], ob, dob)
        return self._result(ob, out), self._result(dob, dout)

    def _out_buffer(self, out, B):
        if out is None:
            return torch.empty(B, self.N, dtype=self.dtype, device=self.device)
        assert isinstance(out, torch.Tensor) and out.shape == (self.N, B) and out.dtype == self.dtype
        assert out.device == self.device, f"out is on {out.device}, the mesh on {self.device}"
        t = out.T
        assert t.is_contiguous(), "out must be an (N, B) Fortran-ordered tensor (transpose of a contiguous (B, N))"
        return t

    @staticmethod
    def _result(buf, out):
        return buf.T if out is None else out

    # ---- loss and gradients (gpu/mesh.py:176-234) -------------------------------------
    def _L2norm_fn(self, J):
        def f(U):
            V = U - torch.eye(self.N, dtype=self.dtype, device=self.device)
            if not self.is_phase:
                V = V - torch.diag_embed(torch.diagonal(V))
            J[0] = float(torch.linalg.norm(V) ** 2)
            return 2 * V
        return f

    def L2norm(self, U_target, p_phase=None, p_splitter=None) -> float:
        """|U_target^H U - 1|^2 over the off-diagonal elements when is_phase is False (gpu/mesh.py:183-193)."""
        J = [0]
        f = self._L2norm_fn(J)
        Ut = torch.as_tensor(np.asarray(U_target) if not isinstance(U_target, torch.Tensor) else U_target)
        f(self.dot(Ut.to(self.dtype).to(self.device).T.conj(), p_phase, p_splitter))
        return J[0]

    def grad_phi_target(self, U_target, p_phase=None, p_splitter=None) -> Tuple[float, torch.Tensor]:
        """(J, dJ/dp) for J = L2norm(U_target) (gpu/mesh.py:195-206)."""
        J = [0]
        f = self._L2norm_fn(J)
        Ut = torch.as_tensor(np.asarray(U_target) if not isinstance(U_target, torch.Tensor) else U_target)
        dp = self.grad_phi(Ut.to(self.dtype).to(self.device).T.conj(), f, p_phase, p_splitter)[0]
        return J[0], dp

    def grad_phi(self, v, dJdv, p_phase=None, p_splitter=None, out_dp=None, out_dJdv=None, vpos="in"):
        """Backpropagation (gpu/mesh.py:208-234).

        vpos='in': v is the input and is propagated forward first; vpos='out': v is the output.
        dJdv: array dJ/dv_out (meshes convention dJ/dRe + i dJ/dIm) or a callable f(v_out) -> dJdv.
        Returns (out_dp, out_dJdv); out_dp (flat, size n_phase) is accumulated into when given.
        """
        assert vpos in ["in", "out"]
        p, s = self._defaults(p_phase, p_splitter)
        if vpos == "in":
            vout = self.dot(v, p_phase, p_splitter)
        else:
            vout = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
        vb = as_field(vout, self.N, self.dtype, self.device)
        B = vb.shape[0]
        if callable(dJdv):
            dJdv = dJdv(vb.T)
        gb = as_field(dJdv, self.N, self.dtype, self.device)
        if out_dp is None:
            out_dp = torch.zeros(self.n_phase, dtype=torch.float32, device=self.device)
        assert isinstance(out_dp, torch.Tensor) and out_dp.dtype == torch.float32 and out_dp.is_contiguous()
        assert out_dp.numel() == self.n_phase
        assert out_dp.device == self.device, f"out_dp is on {out_dp.device}, the mesh on {self.device}"
        gin = self._out_buffer(out_dJdv, B)
        self._backdiff(vb, gb, p, s, out_dp, gin)
        return out_dp, self._result(gin, out_dJdv)

    def _backdiff(self, vb, gb, p, s, out_dp, gin):
        """Adjoint walk of (B, N) output field vb and gradient gb (screens included): accumulates
        dJ/dp into out_dp, writes dJ/dv_in into gin and returns the reconstructed input field."""
        B = vb.shape[0]
        ldp, lds = self._ld()
        g0, g1 = self._geom()
        uin = torch.empty_like(vb)
        n_cr_p = self.n_cr * self.X_np
        if self.is_phase and self.phi_pos == "out":
            # back through the output screen first (on copies: the caller's arrays stay intact)
            vb, gb = vb.clone(), gb.clone()
            self._screen(_lib.OP_BACKDIFF, B, p[n_cr_p:], out_dp[n_cr_p:], vb, gb)
        ws, nb = self._workspace(_lib.OP_BACKDIFF, B)
        _lib.run(_lib.OP_BACKDIFF, self.X, self.N, self.L, B, self.fft, g0, g1, p, out_dp, ldp,
                 s if s.numel() else None, lds, vb, gb, uin, gin, self.N, tile=self.tile, workspace=ws, ws_bytes=nb,
                 flags=self._adj_flags(s))
        if self.is_phase and self.phi_pos == "in":
            self._screen(_lib.OP_BACKDIFF, B, p[n_cr_p:], out_dp[n_cr_p:], uin, gin)
        return uin

    def grad_target_l2(self, v, y, p_phase=None, p_splitter=None, out_dp=None):
        """Fused fwd+grad of J = sum |U v - y|^2 (the loss of gpu/test.py:131): returns
        (J, dp, u_out).  dJ/du_out = 2 (u_out - y) is formed inside the adjoint kernel.  For
        unitary crossings (and generic ones with balanced arm gains) this is one library call (one
        prologue for all tables, fwdprop, adjoint, dp reduction); other lossy crossings run
        fwdprop, then backdiff.  dp is accumulated into out_dp when it is given, else returned
        fresh.

        With a phase screen S = diag(exp(i phi)) (is_phase): at the output, J = |S M v - y|^2 =
        |M v - S^H y|^2, so the mesh part runs against the target S^H y and dJ/dphi comes from the
        screen's adjoint on (u_out, 2 (u_out - y)); at the input, the mesh part runs on S v, returns
        its input gradient, and the screen's adjoint on (S v, that gradient) gives dJ/dphi."""
        p, s = self._defaults(p_phase, p_splitter)
        ldp, lds = self._ld()
        g0, g1 = self._geom()
        vb = as_field(v, self.N, self.dtype, self.device)
        yb = as_field(y, self.N, self.dtype, self.device)
        B = vb.shape[0]
        uout = torch.empty_like(vb)
        flags = 0
        if out_dp is None:
            out_dp = torch.empty(self.n_phase, dtype=torch.float32, device=self.device)
            flags = _lib.DP_ASSIGN
        n_m = self.n_cr * self.X_np
        if self.is_phase and flags:
            out_dp[n_m:].zero_()  # the screen's adjoint adds into its part
        src, tgt, gin = vb, yb, None
        if self.is_phase and self.phi_pos == "in":
            src = vb.clone()
            self._screen(_lib.OP_FWDPROP, B, p[n_m:], None, src, None)
            gin = torch.empty_like(vb)
        elif self.is_phase:
            tgt = yb.clone()
            self._screen(_lib.OP_FWDPROP, B, -p[n_m:], None, tgt, None)  # S^H y
        sp = s if s.numel() else None
        sc = self._adj_flags(s)
        if self.X != "mzi_generic" or sc:
            ws, nb = self._workspace(_lib.OP_FWDGRAD, B)
            _lib.run(_lib.OP_FWDGRAD, self.X, self.N, self.L, B, self.fft, g0, g1, p, out_dp, ldp, sp, lds, src,
                     None, uout, gin, self.N, target=tgt, tile=self.tile, workspace=ws, ws_bytes=nb,
                     flags=flags | sc)
        else:
            ws, nb = self._workspace(_lib.OP_FWDPROP, B)
            _lib.run(_lib.OP_FWDPROP, self.X, self.N, self.L, B, self.fft, g0, g1, p, None, ldp, sp, lds, src, None,
                     uout, None, self.N, tile=self.tile, workspace=ws, ws_bytes=nb)
            uin = torch.empty_like(vb)
            ws, nb = self._workspace(_lib.OP_BACKDIFF, B)
            _lib.run(_lib.OP_BACKDIFF, self.X, self.N, self.L, B, self.fft, g0, g1, p, out_dp, ldp, sp, lds, uout,
                     None, uin, gin, self.N, target=tgt, tile=self.tile, workspace=ws, ws_bytes=nb, flags=flags)
        if self.is_phase and self.phi_pos == "in":
            self._screen(_lib.OP_BACKDIFF, B, p[n_m:], out_dp[n_m:], src, gin)
        elif self.is_phase:
            self._screen(_lib.OP_FWDPROP, B, p[n_m:], None, uout, None)  # u_out = S (M v)
            g = 2 * (uout - yb)
            self._screen(_lib.OP_BACKDIFF, B, p[n_m:], out_dp[n_m:], uout.clone(), g)
        J = torch.sum(torch.abs(uout - yb) ** 2)
        return J, out_dp, uout.T

    # ---- Hessian-vector products (gpu/mesh.py:236-288) --------------------------------
    def exact_hvp_supported(self) -> bool:
        """Forward-over-reverse HVP kernel: unitary crossings (MZI, SYM; orth on NN meshes) up to
        N = 1024 on NN meshes and N = 512 on FFT meshes, with or without a phase screen."""
        if self.X not in ("mzi", "sym", "orth"):
            return False
        if self.fft:
            return self.X != "orth" and self.N <= 512
        return self.N <= 1024

    def _hvp_exact(self, q, out, dout, g, dg, p, s, v=None):
        """H q from the (B, N) output field, its tangent along q, the output gradient and its tangent.

        With an output screen y = S z, S = diag(exp(i phi)): the screen's gradient entries
        sum_b Im(g conj(y)) have the derivative sum_b Im(dg conj(y) + g conj(dy)); the mesh part runs
        on z = S^H y with tangent S^H (dy - i dphi y), gradient S^H g and its tangent
        S^H (dg - i dphi g), as the screen's adjoint hands them to the mesh.  With an input screen
        x = S v: the kernel also returns the input gradient g_x and its tangent dg_x, and the screen's
        entries are sum_b Im(dg_x conj(x) + g_x conj(dx)), dx = i dphi x."""
        B = out.shape[0]
        hv = torch.zeros(self.n_phase, dtype=torch.float32, device=self.device)
        hv_screen = gx = dgx = None
        n_m = self.n_cr * self.X_np
        if self.is_phase and self.phi_pos == "in":
            assert v is not None, "an input screen needs the input field"
            gx, dgx = torch.empty_like(g), torch.empty_like(g)
        elif self.is_phase:
            hv_screen = torch.sum(torch.imag(dg * out.conj() + g * dout.conj()), dim=0)
            sh = torch.exp(-1j * p[n_m:]).to(self.dtype)
            idq = (1j * q[n_m:]).to(self.dtype)
            out, dout = sh * out, sh * (dout - idq * out)
            g, dg = sh * g, sh * (dg - idq * g)
        nb = _lib.hvp_workspace_size(self.X, self.N, self.L, B, self.fft)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.device)
        ldp, lds = self._ld()
        g0, g1 = self._geom()
        _lib.hvp(self.X, self.N, self.L, B, self.fft, g0, g1, p, q, ldp, s if s.numel() else None, lds, out, dout,
                 g, dg, None, None, hv, self.N, ws, nb, dJdu_in=gx, d_dJdu_in=dgx)
        if gx is not None:
            x = v * torch.exp(1j * p[n_m:]).to(self.dtype)
            dx = (1j * q[n_m:]).to(self.dtype) * x
            hv_screen = torch.sum(torch.imag(dgx * x.conj() + gx * dx.conj()), dim=0)
        if hv_screen is not None:
            hv[n_m:] = hv_screen
        return hv

    def _hvp_difference(self, vec, v, w, p, p_splitter, eps):
        """Central difference of grad_phi along vec with step eps (gpu/mesh.py:236-251)."""
        return (self.grad_phi(v, w, p + vec * eps, p_splitter)[0]
                - self.grad_phi(v, w, p - vec * eps, p_splitter)[0]) / (2 * eps)

    def hvp(self, vec, v, w: Callable, p_phase=None, p_splitter=None, eps=1e-3):
        """H vec for the objective whose output gradient is w(v_out).

        Where exact_hvp_supported(): forward-over-reverse on the GPU (fwddiff along vec, then the
        HVP adjoint kernel), with the tangent of w along the output tangent from torch.func.jvp.
        Otherwise -- and when w cannot be traced by torch.func.jvp (it leaves torch, e.g. through
        NumPy or .item(), so its tangent is unknown) -- central differences of grad_phi with step
        eps, as the reference does (gpu/mesh.py:236-251); that is accurate for any w."""
        p = self.p_phase if p_phase is None else as_params(p_phase, self.n_phase, self.device)
        vec = as_params(vec, self.n_phase, self.device)
        if not self.exact_hvp_supported():
            return self._hvp_difference(vec, v, w, p, p_splitter, eps)
        _, s = self._defaults(p, p_splitter)
        out, dout = self.dot(v, p, p_splitter, dp=vec)          # (N, B) views
        try:
            g, dg = torch.func.jvp(w, (out,), (dout,))
        except (RuntimeError, TypeError, ValueError, NotImplementedError):
            return self._hvp_difference(vec, v, w, p, p_splitter, eps)
        gb = as_field(g, self.N, self.dtype, self.device)
        dgb = as_field(dg, self.N, self.dtype, self.device)
        return self._hvp_exact(vec, out.T, dout.T, gb, dgb, p, s, as_field(v, self.N, self.dtype, self.device))

    def hvp_target(self, vec, U_target, p_phase=None, p_splitter=None, eps=1e-3):
        """H vec of L2norm(U_target) (gpu/mesh.py:253-266); exact where supported: the output
        gradient 2 offdiag(U - 1) (2 (U - 1) with a phase screen) is affine, its tangent is
        2 offdiag(dU) (2 dU)."""
        J = [0]
        w = self._L2norm_fn(J)
        Ut = torch.as_tensor(np.asarray(U_target) if not isinstance(U_target, torch.Tensor) else U_target)
        v = Ut.to(self.dtype).to(self.device).T.conj()
        if not self.exact_hvp_supported():
            return self.hvp(vec, v, w, p_phase, p_splitter, eps)
        p = self.p_phase if p_phase is None else as_params(p_phase, self.n_phase, self.device)
        q = as_params(vec, self.n_phase, self.device)
        _, s = self._defaults(p, p_splitter)
        out, dout = self.dot(v, p, p_splitter, dp=q)
        g = w(out)
        dg = 2 * dout
        if not self.is_phase:  # the loss's diagonal is dropped without a screen
            dg = dg - torch.diag_embed(torch.diagonal(dg))
        return self._hvp_exact(q, out.T, dout.T, as_field(g, self.N, self.dtype, self.device),
                               as_field(dg, self.N, self.dtype, self.device), p, s,
                               as_field(v, self.N, self.dtype, self.device))

    def hess(self, v, w: Callable, p_phase=None, p_splitter=None, eps=1e-3):
        # (the reference passes these arguments to hvp_target by mistake, gpu/mesh.py:276)
        vec = torch.eye(self.n_phase, device=self.device)
        return torch.stack([self.hvp(vec_i, v, w, p_phase, p_splitter, eps) for vec_i in vec])

    def hess_target(self, U_target, p_phase=None, p_splitter=None, eps=1e-3):
        vec = torch.eye(self.n_phase, device=self.device)
        return torch.stack([self.hvp_target(vec_i, U_target, p_phase, p_splitter, eps) for vec_i in vec])

    def copy(self) -> "MeshNetworkGPU":
        g = dict(strides=self.strides.cpu().numpy()) if self.fft else {}
        return MeshNetworkGPU(self.N, self.L, self.lens.cpu().numpy(), self.shifts.cpu().numpy(),
                              self.p_phase.clone(), self.p_splitter.clone(), self.X, self.phi_pos, self.is_phase,
                              tile=self.tile, **g)
