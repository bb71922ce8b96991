"""FP64 NumPy oracle for batched field propagation through meshes of 2x2 crossings.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``meshes_amd``) may import this
module; it is the checker that the HIP path is compared against, and it is what
``bench.py`` times as the ``port`` CPU baseline when the compiled reference engine
is not available.

Every function restates the reference algorithm in double precision and cites the
reference file:line it follows.  Conventions (all shared with the GPU path):

* Parameters use the *slot layout* of the reference GPU code: ``p[L, N//2, n_p]``
  and ``s[L, N//2, n_s]``; slot ``m`` of a nearest-neighbour (NN) layer couples
  waveguides ``(2m + shift%2, 2m + shift%2 + 1)`` and is active iff
  ``shift//2 <= m < shift//2 + len``; inactive slots are the identity
  (gpu/gmem.cu:236-240, 264-273; mesh.py:513-522).
* FFT / butterfly layers with stride index ``st`` (stride ``S = 2**st``): slot
  ``c = blk*S + j`` couples ``(2*S*blk + j, 2*S*blk + j + S)`` and every slot is
  active (gpu/test.py:73-77; gpu/fwdprop.cu:48-93).
* Fields are ``(B, N)`` arrays, one vector per row -- the transpose view of the
  reference's Fortran ``(N, B)`` arrays (gpu/mesh.py:150-153).
* Gradients follow the meshes convention ``g = dJ/dRe(u) + i dJ/dIm(u)``
  (``2 dJ/du*``) and ``dJ/dp = Re sum_ij dJdT_ij dT_ij/dp`` with
  ``dJdT_ij = sum_b conj(g_i) x_j`` (crossing.py:146-163; c/dot.cpp:240-278).

Parity status of each function (what pins it) is listed in ``PINS`` at the end.
"""
from __future__ import annotations

import numpy as np

# Number of phase / splitter parameters per crossing type (gpu/mesh.py:73-75;
# crossing.py:444-455 for the generic lossy MZI, which has 14 splitter values).
N_PHASE = {"mzi": 2, "sym": 2, "orth": 1, "mzi_generic": 2}
N_SPLIT = {"mzi": 2, "sym": 1, "orth": 0, "mzi_generic": 14}
REAL_FIELD = {"mzi": False, "sym": False, "orth": True, "mzi_generic": False}


# --------------------------------------------------------------------------------------
# Crossing transfer matrices
# --------------------------------------------------------------------------------------
def _split(p, s, X):
    p = np.asarray(p, dtype=np.float64)
    n_s = N_SPLIT[X]
    if s is None or np.size(s) == 0:
        s = np.zeros(p.shape[:-1] + (n_s,))
    s = np.broadcast_to(np.asarray(s, dtype=np.float64), p.shape[:-1] + (n_s,))
    return p, s


def crossing_T(X, p, s=None):
    """Transfer matrix T[..., 2, 2] (complex128) of each crossing.

    mzi   : crossing.py:216-221 / gpu/crossing.cu:36-62
    sym   : crossing.py:377-381 / gpu/crossing.cu:86-106
    orth  : gpu/crossing.cu:128-139 (real; returned as complex with zero imaginary part)
    mzi_generic : crossing.py:465-488 (MZICrossingGeneric, out_phase=False)
    """
    p, s = _split(p, s, X)
    if X == "mzi":
        th, ph = p[..., 0], p[..., 1]
        a, b = s[..., 0], s[..., 1]
        C, S = np.cos(th / 2), np.sin(th / 2)
        Cp, Sp, Cm, Sm = np.cos(a + b), np.sin(a + b), np.cos(a - b), np.sin(a - b)
        g = np.exp(1j * th / 2)
        e = np.exp(1j * ph)
        T = np.empty(p.shape[:-1] + (2, 2), dtype=np.complex128)
        T[..., 0, 0] = g * e * (1j * S * Cm - C * Sp)
        T[..., 0, 1] = g * (1j * C * Cp - S * Sm)
        T[..., 1, 0] = g * e * (1j * C * Cp + S * Sm)
        T[..., 1, 1] = g * (-1j * S * Cm - C * Sp)
        return T
    if X == "sym":
        th, ph = p[..., 0], p[..., 1]
        al = s[..., 0]
        C, S = np.cos(th / 2), np.sin(th / 2)
        C2, S2 = np.cos(2 * al), np.sin(2 * al)
        T = np.empty(p.shape[:-1] + (2, 2), dtype=np.complex128)
        T[..., 0, 0] = np.exp(1j * ph) * (S + 1j * C * S2)
        T[..., 0, 1] = 1j * C * C2
        T[..., 1, 0] = 1j * C * C2
        T[..., 1, 1] = np.exp(-1j * ph) * (S - 1j * C * S2)
        return T
    if X == "orth":
        th = p[..., 0]
        C, S = np.cos(th / 2), np.sin(th / 2)
        T = np.empty(p.shape[:-1] + (2, 2), dtype=np.complex128)
        T[..., 0, 0] = S
        T[..., 0, 1] = -C
        T[..., 1, 0] = C
        T[..., 1, 1] = S
        return T
    if X == "mzi_generic":
        return _generic_parts(p, s)[0]
    raise ValueError(X)


def _generic_parts(p, s):
    """T4 @ T3 @ T2 @ T1 of the generic lossy MZI (crossing.py:465-488).

    Returns (T, T4, T3, T2, T1) with shapes (..., 2, 2)."""
    th, ph = p[..., 0], p[..., 1]
    (a, b, p11, p21, p12, p22, q11, q21, q12, q22, g11, g21, g12, g22) = [s[..., k] for k in range(14)]
    z = np.zeros_like(th)
    sh = th.shape + (2, 2)
    T1 = np.zeros(sh, np.complex128)
    T1[..., 0, 0] = np.exp(1j * (ph + p11 + q11) + g11 / 2)
    T1[..., 1, 1] = np.exp(1j * (p21 + q21) + g21 / 2)
    T3 = np.zeros(sh, np.complex128)
    T3[..., 0, 0] = np.exp(1j * (th + p12 + q12) + g12 / 2)
    T3[..., 1, 1] = np.exp(1j * (p22 + q22) + g22 / 2)

    def split(x):
        M = np.empty(sh, np.complex128)
        M[..., 0, 0] = np.cos(np.pi / 4 + x) + z
        M[..., 0, 1] = 1j * np.sin(np.pi / 4 + x) + z
        M[..., 1, 0] = 1j * np.sin(np.pi / 4 + x) + z
        M[..., 1, 1] = np.cos(np.pi / 4 + x) + z
        return M

    T2, T4 = split(a), split(b)
    T = T4 @ (T3 @ (T2 @ T1))
    return T, T4, T3, T2, T1


def crossing_dT(X, p, s=None):
    """dT/dp_k, shape (..., n_p, 2, 2).

    mzi: closed form of crossing.py:223-231 (written here via the product rule
    dT/dθ = (i/2) T + e^{iθ/2} dM/dθ, dT/dφ = i T[:,0] on column 0).
    sym / orth: derivatives of the closed forms in gpu/crossing.cu:104-114, 138-144.
    mzi_generic: dT/dφ = T·diag(i,0); dT/dθ = T4·diag(i·T3_00, 0)·T2·T1 (product rule on
    crossing.py:465-488).  All four are pinned by central finite differences in the tests.
    """
    p, s = _split(p, s, X)
    out = np.zeros(p.shape[:-1] + (N_PHASE[X], 2, 2), dtype=np.complex128)
    if X == "mzi":
        th, ph = p[..., 0], p[..., 1]
        a, b = s[..., 0], s[..., 1]
        C, S = np.cos(th / 2), np.sin(th / 2)
        Cp, Sp, Cm, Sm = np.cos(a + b), np.sin(a + b), np.cos(a - b), np.sin(a - b)
        g = np.exp(1j * th / 2)
        e = np.exp(1j * ph)
        T = crossing_T(X, p, s)
        # dM/dθ where T = g * M, S' = C/2, C' = -S/2
        dM = np.empty_like(T)
        dM[..., 0, 0] = e * (1j * (C / 2) * Cm + (S / 2) * Sp)
        dM[..., 0, 1] = (1j * (-S / 2) * Cp - (C / 2) * Sm)
        dM[..., 1, 0] = e * (1j * (-S / 2) * Cp + (C / 2) * Sm)
        dM[..., 1, 1] = (-1j * (C / 2) * Cm + (S / 2) * Sp)
        out[..., 0, :, :] = 0.5j * T + g[..., None, None] * dM
        out[..., 1, :, 0] = 1j * T[..., :, 0]
        return out
    if X == "sym":
        th, ph = p[..., 0], p[..., 1]
        al = s[..., 0]
        C, S = np.cos(th / 2), np.sin(th / 2)
        C2, S2 = np.cos(2 * al), np.sin(2 * al)
        out[..., 0, 0, 0] = np.exp(1j * ph) * (C / 2 - 1j * (S / 2) * S2)
        out[..., 0, 0, 1] = -1j * (S / 2) * C2
        out[..., 0, 1, 0] = -1j * (S / 2) * C2
        out[..., 0, 1, 1] = np.exp(-1j * ph) * (C / 2 + 1j * (S / 2) * S2)
        T = crossing_T(X, p, s)
        out[..., 1, 0, 0] = 1j * T[..., 0, 0]
        out[..., 1, 1, 1] = -1j * T[..., 1, 1]
        return out
    if X == "orth":
        th = p[..., 0]
        C, S = np.cos(th / 2), np.sin(th / 2)
        out[..., 0, 0, 0] = C / 2
        out[..., 0, 0, 1] = S / 2
        out[..., 0, 1, 0] = -S / 2
        out[..., 0, 1, 1] = C / 2
        return out
    if X == "mzi_generic":
        T, T4, T3, T2, T1 = _generic_parts(p, s)
        D = np.zeros_like(T)
        D[..., 0, 0] = 1j
        out[..., 1, :, :] = T @ D
        dT3 = np.zeros_like(T3)
        dT3[..., 0, 0] = 1j * T3[..., 0, 0]
        out[..., 0, :, :] = T4 @ (dT3 @ (T2 @ T1))
        return out
    raise ValueError(X)


def crossing_d2T(X, p, s=None):
    """Second derivatives d2T/dp_k dp_m, shape (..., n_p, n_p, 2, 2), of the closed forms.

    Derived from crossing_T / crossing_dT (the derivatives only exist here and on the GPU;
    the reference differentiates once).  mzi and mzi_generic: every theta-dependent factor is
    a phase or linear in (cos th/2, sin th/2) times a phase, which gives d2T/dth2 = i dT/dth,
    d2T/dth dph = dT/dth diag(i, 0), d2T/dph2 = T diag(-1, 0).  sym: entries are linear in
    (cos th/2, sin th/2): d2T/dth2 = -T/4; phi enters as exp(+-i ph) on the diagonal.
    orth: -T/4.  Pinned by central finite differences of crossing_dT."""
    p, s = _split(p, s, X)
    T = crossing_T(X, p, s)
    dT = crossing_dT(X, p, s)
    n_p = N_PHASE[X]
    out = np.zeros(p.shape[:-1] + (n_p, n_p, 2, 2), dtype=np.complex128)
    if X in ("mzi", "mzi_generic"):
        out[..., 0, 0, :, :] = 1j * dT[..., 0, :, :]
        out[..., 0, 1, :, 0] = 1j * dT[..., 0, :, 0]
        out[..., 1, 0, :, :] = out[..., 0, 1, :, :]
        out[..., 1, 1, :, 0] = -T[..., :, 0]
    elif X == "sym":
        out[..., 0, 0, :, :] = -0.25 * T
        out[..., 0, 1, 0, 0] = 1j * dT[..., 0, 0, 0]
        out[..., 0, 1, 1, 1] = -1j * dT[..., 0, 1, 1]
        out[..., 1, 0, :, :] = out[..., 0, 1, :, :]
        out[..., 1, 1, 0, 0] = -T[..., 0, 0]
        out[..., 1, 1, 1, 1] = -T[..., 1, 1]
    elif X == "orth":
        out[..., 0, 0, :, :] = -0.25 * T
    return out


def hvp_inverse(geom, X, p, s, u_in, q, target):
    """Exact Hessian-vector product of J = sum |U(p) u_in - target|^2 along q by
    forward-over-reverse (unitary crossings): the forward JVP gives (u_out, du_out); the
    reverse walk carries (u, du, g, dg) with x = T^H u, dx = T^H du + dT_q^H u,
    g <- T^H g, dg <- T^H dg + dT_q^H g (dT_q = sum_k q_k dT/dp_k), accumulates
    dJdT = sum conj(g) x and its tangent sum conj(dg) x + conj(g) dx, and returns
    (grad, hvp) with hvp_k = Re sum_ij [d(dJdT)_ij dT_ij/dp_k + dJdT_ij sum_m q_m d2T_ij/dp_k dp_m].
    Inactive slots give zero.  Pinned by central differences of the FP64 gradient along q."""
    T = slot_T(geom, X, p, s)
    n_p = N_PHASE[X]
    pp = np.asarray(p, dtype=np.float64).reshape(geom.L, geom.N // 2, n_p)
    ss = None if (s is None or np.size(s) == 0) else np.asarray(s, np.float64).reshape(geom.L, geom.N // 2, N_SPLIT[X])
    qq = np.asarray(q, dtype=np.float64).reshape(pp.shape)
    dTk = crossing_dT(X, pp, ss)                                  # [L, N/2, n_p, 2, 2]
    dTq = np.einsum("lmkij,lmk->lmij", dTk, qq)
    u, du = jvp(geom, T, dTq, u_in)
    g = 2 * (u - np.asarray(target, np.complex128))
    dg = 2 * du
    dJdT = np.zeros(T.shape, dtype=np.complex128)
    ddJdT = np.zeros(T.shape, dtype=np.complex128)
    for l in range(geom.L - 1, -1, -1):
        m, i, j = geom.pairs(l)
        th = np.conj(np.swapaxes(T[l, m], -1, -2))
        dth = np.conj(np.swapaxes(dTq[l, m], -1, -2))

        def ap(M, a, b):
            return M[:, 0, 0] * a + M[:, 0, 1] * b, M[:, 1, 0] * a + M[:, 1, 1] * b
        ui, uj, dui, duj = u[:, i].copy(), u[:, j].copy(), du[:, i].copy(), du[:, j].copy()
        gi, gj, dgi, dgj = g[:, i].copy(), g[:, j].copy(), dg[:, i].copy(), dg[:, j].copy()
        xi, xj = ap(th, ui, uj)
        a1, a2 = ap(th, dui, duj)
        b1, b2 = ap(dth, ui, uj)
        dxi, dxj = a1 + b1, a2 + b2
        for (r, gr, dgr) in ((0, gi, dgi), (1, gj, dgj)):
            for (c, xc, dxc) in ((0, xi, dxi), (1, xj, dxj)):
                dJdT[l, m, r, c] = np.sum(np.conj(gr) * xc, axis=0)
                ddJdT[l, m, r, c] = np.sum(np.conj(dgr) * xc + np.conj(gr) * dxc, axis=0)
        u[:, i], u[:, j] = xi, xj
        du[:, i], du[:, j] = dxi, dxj
        g[:, i], g[:, j] = ap(th, gi, gj)
        c1, c2 = ap(th, dgi, dgj)
        e1, e2 = ap(dth, gi, gj)
        dg[:, i], dg[:, j] = c1 + e1, c2 + e2
    d2 = crossing_d2T(X, pp, ss)                                  # [L, N/2, n_p, n_p, 2, 2]
    Sq = np.einsum("lmkaij,lma->lmkij", d2, qq)
    grad = np.real(np.einsum("lmkij,lmij->lmk", dTk, dJdT))
    hv = np.real(np.einsum("lmkij,lmij->lmk", dTk, ddJdT) + np.einsum("lmkij,lmij->lmk", Sq, dJdT))
    act = active_mask(geom)
    grad[~act] = 0.0
    hv[~act] = 0.0
    return grad, hv


def grad_from_dJdT(X, p, s, dJdT):
    """dJ/dp_k = Re sum_ij dJdT_ij dT_ij/dp_k (crossing.py:146-163, gpu/crossing.cu:150-203)."""
    dT = crossing_dT(X, p, s)
    return np.real(np.einsum("...kij,...ij->...k", dT, dJdT))


# --------------------------------------------------------------------------------------
# Geometry
# --------------------------------------------------------------------------------------
class Geometry:
    """Layer geometry of a mesh: NN (lens/shifts) or FFT (strides).

    NN semantics: c/dot.cpp:148-163 and gpu/gmem.cu:236-240; FFT: gpu/test.py:73-77.
    """

    def __init__(self, N, L, lens=None, shifts=None, strides=None):
        self.N, self.L = int(N), int(L)
        assert self.N % 2 == 0
        if strides is not None:
            self.kind = "fft"
            self.strides = np.asarray(strides, dtype=np.int64).reshape(L)
            for st in self.strides:
                assert 0 <= st and self.N % (2 << int(st)) == 0, "stride does not tile N"
            self.lens = np.full(L, self.N // 2, dtype=np.int64)
            self.shifts = np.zeros(L, dtype=np.int64)
        else:
            self.kind = "nn"
            self.lens = np.asarray(lens, dtype=np.int64).reshape(L)
            self.shifts = np.asarray(shifts, dtype=np.int64).reshape(L)
            assert np.all(self.shifts >= 0) and np.all(self.lens >= 0)
            assert np.all(self.shifts + 2 * self.lens <= self.N), "crossing beyond the last waveguide"
            self.strides = None

    def pairs(self, l):
        """(slots m, upper waveguides i, lower waveguides j) of the active crossings of layer l."""
        if self.kind == "fft":
            S = 1 << int(self.strides[l])
            c = np.arange(self.N // 2)
            blk, j = c // S, c % S
            i = 2 * S * blk + j
            return c, i, i + S
        sh, ln = int(self.shifts[l]), int(self.lens[l])
        m = np.arange(sh // 2, sh // 2 + ln)
        i = 2 * m + sh % 2
        return m, i, i + 1

    @property
    def n_active(self):
        return int(np.sum(self.lens))


def clements(N):
    """Clements rectangular mesh: L = N, shifts[l] = l%2, lens[l] = (N - l%2)//2 (clements.py:56-57)."""
    l = np.arange(N)
    return Geometry(N, N, lens=(N - l % 2) // 2, shifts=l % 2)


def butterfly_strides(N):
    """Butterfly / FFT stride indices 0..log2(N)-1 (one layer per stride)."""
    n = int(np.log2(N))
    assert 1 << n == N
    return np.arange(n)


# --------------------------------------------------------------------------------------
# Propagation (fwd / JVP / VJP)
# --------------------------------------------------------------------------------------
def _T_layers(geom, T):
    T = np.asarray(T)
    assert T.shape == (geom.L, geom.N // 2, 2, 2)
    return T


def fwd(geom, T, u):
    """u_out = M_{L-1} ... M_0 u_in for every row of u (c/dot.cpp:146-238 mode 0)."""
    T = _T_layers(geom, T)
    u = np.array(u, dtype=np.complex128, copy=True)
    for l in range(geom.L):
        m, i, j = geom.pairs(l)
        t = T[l, m]
        ui, uj = u[:, i].copy(), u[:, j].copy()
        u[:, i] = t[:, 0, 0] * ui + t[:, 0, 1] * uj
        u[:, j] = t[:, 1, 0] * ui + t[:, 1, 1] * uj
    return u


def jvp(geom, T, dT, u, du=None):
    """(u_out, du_out) with du' = T du + dT u per crossing (c/dot.cpp:189-216; gpu/crossing.cu:240-251).

    ``dT`` is the directional derivative sum_k dp_k dT/dp_k in slot layout (mesh.py:368)."""
    T = _T_layers(geom, T)
    dT = _T_layers(geom, dT)
    u = np.array(u, dtype=np.complex128, copy=True)
    du = np.zeros_like(u) if du is None else np.array(du, dtype=np.complex128, copy=True)
    for l in range(geom.L):
        m, i, j = geom.pairs(l)
        t, d = T[l, m], dT[l, m]
        ui, uj, dui, duj = u[:, i].copy(), u[:, j].copy(), du[:, i].copy(), du[:, j].copy()
        du[:, i] = t[:, 0, 0] * dui + t[:, 0, 1] * duj + d[:, 0, 0] * ui + d[:, 0, 1] * uj
        du[:, j] = t[:, 1, 0] * dui + t[:, 1, 1] * duj + d[:, 1, 0] * ui + d[:, 1, 1] * uj
        u[:, i] = t[:, 0, 0] * ui + t[:, 0, 1] * uj
        u[:, j] = t[:, 1, 0] * ui + t[:, 1, 1] * uj
    return u, du


def vjp_inverse(geom, T, u_out, g_out):
    """Adjoint by inverse propagation, valid for unitary crossings (c/dot.cpp:240-278 mode 2;
    gpu/backdiff.cu:24-153 + gpu/crossing.cu:285-302).

    Walks layers L-1..0: x = T^H u; dJdT_ij += sum_b conj(g_i) x_j; g <- T^H g.
    Returns (u_in reconstructed, g_in, dJdT[L, N/2, 2, 2])."""
    T = _T_layers(geom, T)
    u = np.array(u_out, dtype=np.complex128, copy=True)
    g = np.array(g_out, dtype=np.complex128, copy=True)
    dJdT = np.zeros(T.shape, dtype=np.complex128)
    for l in range(geom.L - 1, -1, -1):
        m, i, j = geom.pairs(l)
        th = np.conj(np.swapaxes(T[l, m], -1, -2))  # T^H
        ui, uj, gi, gj = u[:, i].copy(), u[:, j].copy(), g[:, i].copy(), g[:, j].copy()
        xi = th[:, 0, 0] * ui + th[:, 0, 1] * uj
        xj = th[:, 1, 0] * ui + th[:, 1, 1] * uj
        dJdT[l, m, 0, 0] = np.sum(np.conj(gi) * xi, axis=0)
        dJdT[l, m, 0, 1] = np.sum(np.conj(gi) * xj, axis=0)
        dJdT[l, m, 1, 0] = np.sum(np.conj(gj) * xi, axis=0)
        dJdT[l, m, 1, 1] = np.sum(np.conj(gj) * xj, axis=0)
        u[:, i], u[:, j] = xi, xj
        g[:, i] = th[:, 0, 0] * gi + th[:, 0, 1] * gj
        g[:, j] = th[:, 1, 0] * gi + th[:, 1, 1] * gj
    return u, g, dJdT


def vjp_stored(geom, T, u_in, g_out):
    """Exact adjoint for any (possibly non-unitary) crossings: the forward fields are stored
    layer by layer instead of being reconstructed.  Same accumulation as vjp_inverse, with
    the true layer inputs x and g <- T^H g.  Returns (u_out, g_in, dJdT)."""
    T = _T_layers(geom, T)
    u = np.array(u_in, dtype=np.complex128, copy=True)
    inputs = []
    for l in range(geom.L):
        inputs.append(u.copy())
        m, i, j = geom.pairs(l)
        t = T[l, m]
        ui, uj = u[:, i].copy(), u[:, j].copy()
        u[:, i] = t[:, 0, 0] * ui + t[:, 0, 1] * uj
        u[:, j] = t[:, 1, 0] * ui + t[:, 1, 1] * uj
    u_out = u
    g = np.array(g_out, dtype=np.complex128, copy=True)
    dJdT = np.zeros(T.shape, dtype=np.complex128)
    for l in range(geom.L - 1, -1, -1):
        m, i, j = geom.pairs(l)
        x = inputs[l]
        th = np.conj(np.swapaxes(T[l, m], -1, -2))
        xi, xj, gi, gj = x[:, i], x[:, j], g[:, i].copy(), g[:, j].copy()
        dJdT[l, m, 0, 0] = np.sum(np.conj(gi) * xi, axis=0)
        dJdT[l, m, 0, 1] = np.sum(np.conj(gi) * xj, axis=0)
        dJdT[l, m, 1, 0] = np.sum(np.conj(gj) * xi, axis=0)
        dJdT[l, m, 1, 1] = np.sum(np.conj(gj) * xj, axis=0)
        g[:, i] = th[:, 0, 0] * gi + th[:, 0, 1] * gj
        g[:, j] = th[:, 1, 0] * gi + th[:, 1, 1] * gj
    return u_out, g, dJdT


def slot_T(geom, X, p, s=None):
    """T in slot layout [L, N/2, 2, 2] from slot-layout parameters (gpu/gmem.cu:264-273)."""
    n_p = N_PHASE[X]
    p = np.asarray(p, dtype=np.float64).reshape(geom.L, geom.N // 2, n_p)
    if s is not None and np.size(s):
        s = np.asarray(s, dtype=np.float64).reshape(geom.L, geom.N // 2, N_SPLIT[X])
    else:
        s = None
    return crossing_T(X, p, s)


def slot_dT_dir(geom, X, p, s, dp):
    """Directional derivative sum_k dp_k dT/dp_k in slot layout (mesh.py:368 einsum)."""
    n_p = N_PHASE[X]
    p = np.asarray(p, dtype=np.float64).reshape(geom.L, geom.N // 2, n_p)
    dp = np.asarray(dp, dtype=np.float64).reshape(geom.L, geom.N // 2, n_p)
    if s is not None and np.size(s):
        s = np.asarray(s, dtype=np.float64).reshape(geom.L, geom.N // 2, N_SPLIT[X])
    else:
        s = None
    return np.einsum("lmkij,lmk->lmij", crossing_dT(X, p, s), dp)


def active_mask(geom):
    """Boolean [L, N/2]: which slots hold an active crossing."""
    mask = np.zeros((geom.L, geom.N // 2), dtype=bool)
    for l in range(geom.L):
        mask[l, geom.pairs(l)[0]] = True
    return mask


# --------------------------------------------------------------------------------------
# High-level helpers mirroring gpu/mesh.py semantics
# --------------------------------------------------------------------------------------
def mesh_fwd(geom, X, p, s, u_in):
    return fwd(geom, slot_T(geom, X, p, s), u_in)


def mesh_jvp(geom, X, p, s, u_in, dp=None, du_in=None):
    T = slot_T(geom, X, p, s)
    dT = np.zeros_like(T) if dp is None else slot_dT_dir(geom, X, p, s, dp)
    return jvp(geom, T, dT, u_in, du_in)


def mesh_vjp(geom, X, p, s, u_out, g_out, u_in=None):
    """(dp[L, N/2, n_p], g_in, u_in) of the backdiff kernel.

    With ``u_in`` given, uses stored forward fields (exact for lossy crossings); otherwise
    reconstructs by inverse propagation like the reference (unitary crossings only).
    Inactive slots get zero gradient (the reference accumulates only active slots,
    gpu/gmem.cu:309-318 with cond_gen)."""
    T = slot_T(geom, X, p, s)
    if u_in is None:
        u_rec, g_in, dJdT = vjp_inverse(geom, T, u_out, g_out)
    else:
        _, g_in, dJdT = vjp_stored(geom, T, u_in, g_out)
        u_rec = np.array(u_in, dtype=np.complex128)
    n_p = N_PHASE[X]
    pp = np.asarray(p, dtype=np.float64).reshape(geom.L, geom.N // 2, n_p)
    ss = None if (s is None or np.size(s) == 0) else np.asarray(s, np.float64).reshape(geom.L, geom.N // 2, N_SPLIT[X])
    dp = grad_from_dJdT(X, pp, ss, dJdT)
    dp[~active_mask(geom)] = 0.0
    return dp, g_in, u_rec


def dense_matrix(geom, T):
    """Brute-force dense product of block-diagonal layer matrices (gpu/test.py:73-80)."""
    N = geom.N
    U = np.eye(N, dtype=np.complex128)
    for l in range(geom.L):
        M = np.eye(N, dtype=np.complex128)
        m, i, j = geom.pairs(l)
        t = T[l, m]
        M[i, i], M[i, j], M[j, i], M[j, j] = t[:, 0, 0], t[:, 0, 1], t[:, 1, 0], t[:, 1, 1]
        U = M @ U
    return U


def random_geometry_nn(rng, N, L):
    """Random NN geometry drawn like gpu/test.py:41 (shifts in [0, N-2], lens in [0, (N-shift)/2))."""
    shifts = rng.randint(0, N - 1, size=L)
    lens = np.array([rng.randint(0, max(1, (N - sh) // 2)) for sh in shifts])
    return Geometry(N, L, lens=lens, shifts=shifts)


def unit_cn(rng, B, N, real=False):
    """B i.i.d. CN(0,1) vectors of length N, each normalised to unit norm."""
    if real:
        x = rng.standard_normal((B, N))
    else:
        x = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# --------------------------------------------------------------------------------------
# Packed (StructuredMeshNetwork) form: permutations and the phase screen
# --------------------------------------------------------------------------------------
def _screen(v, ph):
    return v * np.exp(1j * np.asarray(ph, np.float64))[None, :]


def structured_run(N, lens, shifts, T, v, mode, dT=None, dv=None, ph=None, dph=None, perms=None, pos=-1):
    """The reference CPU engine's three modes on the packed layout (c/dot.cpp:49-308,
    mesh.py:29-70): crossings of layer n are T[:, :, inds[n]:inds[n]+lens[n]] acting on
    waveguides (shifts[n] + 2k, shifts[n] + 2k + 1); ``perms[n]`` (gather, v <- v[p]) is applied
    before layer n and ``perms[L]`` after the last; the phase screen exp(i ph) sits at the input
    (pos 0) or the output (pos 1), none for pos -1.  Fields are (B, N) rows.

    mode 0: returns v_out.  mode 1: returns (v_out, dv_out) with the tangent T' = dT, screen
    tangent dph.  mode 2 (v = v_out, dv = dJ/dv_out, ``perms`` the inverse permutations as
    mesh.py:381 passes them): returns (v_in, dJ/dv_in, dJdT (2, 2, nT), dJ/dph)."""
    T = np.asarray(T, np.complex128)
    v = np.array(v, dtype=np.complex128, copy=True)
    L = len(lens)
    inds = np.cumsum([0] + list(lens))
    is_fd, is_bd = mode == 1, mode == 2
    dv = (np.zeros_like(v) if dv is None else np.array(dv, dtype=np.complex128, copy=True)) if (is_fd or is_bd) else None
    ph = np.zeros(N) if ph is None else np.asarray(ph, np.float64)
    dph = np.zeros(N) if dph is None else np.asarray(dph, np.float64)
    dph_out = np.zeros(N)
    dJdT = np.zeros_like(T) if is_bd else None
    is_perm = perms is not None and len(perms) > 0
    # c/dot.cpp:82-87: the adjoint walks backwards with T^H (conjugate, transposed
    # off-diagonals) and the screen at the mirrored position with -ph.
    if is_bd:
        pos = 1 - pos if pos >= 0 else pos
        Tw = np.conj(T).transpose(1, 0, 2)
        phw = -ph
        order = range(L - 1, -1, -1)
    else:
        Tw = T
        phw = ph
        order = range(L)

    def screen_step():
        nonlocal v, dv
        v = _screen(v, phw)
        if is_fd:
            dv = _screen(dv, phw) + 1j * v * dph[None, :]
        if is_bd:
            dv = _screen(dv, phw)
            dph_out[:] += np.sum(v.real * dv.imag - v.imag * dv.real, axis=0)

    if pos == 0:
        screen_step()
    for n in order:
        if is_perm:
            p = np.asarray(perms[n + int(is_bd)])
            v = v[:, p]
            if is_fd or is_bd:
                dv = dv[:, p]
        i1, ln, s = inds[n], lens[n], shifts[n]
        a, b = np.arange(s, s + 2 * ln, 2), np.arange(s + 1, s + 2 * ln, 2)
        t = Tw[:, :, i1:i1 + ln]
        v1, v2 = v[:, a].copy(), v[:, b].copy()
        if is_fd:
            d = dT[:, :, i1:i1 + ln]
            d1, d2 = dv[:, a].copy(), dv[:, b].copy()
            dv[:, a] = d[0, 0] * v1 + d[0, 1] * v2 + t[0, 0] * d1 + t[0, 1] * d2
            dv[:, b] = d[1, 0] * v1 + d[1, 1] * v2 + t[1, 0] * d1 + t[1, 1] * d2
        v[:, a] = t[0, 0] * v1 + t[0, 1] * v2
        v[:, b] = t[1, 0] * v1 + t[1, 1] * v2
        if is_bd:
            # dJ/dT_ij = sum_b conj(dJ/dy_i) x_j with x the reconstructed layer input
            # (c/dot.cpp:240-285)
            g1, g2 = dv[:, a].copy(), dv[:, b].copy()
            x1, x2 = v[:, a], v[:, b]
            dJdT[0, 0, i1:i1 + ln] += np.sum(np.conj(g1) * x1, axis=0)
            dJdT[0, 1, i1:i1 + ln] += np.sum(np.conj(g1) * x2, axis=0)
            dJdT[1, 0, i1:i1 + ln] += np.sum(np.conj(g2) * x1, axis=0)
            dJdT[1, 1, i1:i1 + ln] += np.sum(np.conj(g2) * x2, axis=0)
            dv[:, a] = t[0, 0] * g1 + t[0, 1] * g2
            dv[:, b] = t[1, 0] * g1 + t[1, 1] * g2
    if is_perm:
        p = np.asarray(perms[0 if is_bd else L])
        v = v[:, p]
        if is_fd or is_bd:
            dv = dv[:, p]
    if pos == 1:
        screen_step()
    if mode == 0:
        return v
    if mode == 1:
        return v, dv
    return v, dv, dJdT, dph_out


def structured_dense(N, lens, shifts, T, perms=None, ph=None, pos=-1):
    """Dense FP64 transfer matrix of the packed form (brute force for small N):
    U = S_out P_L M_{L-1} P_{L-1} ... M_0 P_0 S_in."""
    T = np.asarray(T, np.complex128)
    inds = np.cumsum([0] + list(lens))
    U = np.eye(N, dtype=np.complex128)
    S = np.diag(np.exp(1j * np.asarray(ph, np.float64))) if (ph is not None and pos >= 0) else np.eye(N)
    if pos == 0:
        U = S @ U
    for n in range(len(lens)):
        if perms is not None and len(perms):
            U = np.eye(N)[np.asarray(perms[n])] @ U
        M = np.eye(N, dtype=np.complex128)
        for k in range(lens[n]):
            i, j = shifts[n] + 2 * k, shifts[n] + 2 * k + 1
            M[np.ix_([i, j], [i, j])] = T[:, :, inds[n] + k]
        U = M @ U
    if perms is not None and len(perms):
        U = np.eye(N)[np.asarray(perms[len(lens)])] @ U
    if pos == 1:
        U = S @ U
    return U


def butterfly_perms(N):
    """ButterflyNetwork geometry (butterfly.py:46-51): N-1 full layers; before layer i (odd)
    the gather that makes stride-s pairs adjacent, s = 2^(trailing zeros of i+1), and before
    layer i+1 its inverse."""
    lens, shifts, perm = [N // 2] * (N - 1), [0] * (N - 1), [None] * N
    for i in range(1, N - 1, 2):
        s = 1 << ((i + 1) & -(i + 1)).bit_length() - 1
        blocks = np.arange(0, N, 2 * s)
        p1 = (blocks[:, None] + (np.arange(s)[:, None] + np.array([[0, s]])).reshape(-1)[None, :]).reshape(-1)
        p2 = (blocks[:, None] + (np.arange(0, 2 * s, 2)[None, :] + np.array([[0], [1]])).reshape(-1)[None, :]).reshape(-1)
        perm[i], perm[i + 1] = p1, p2
    full = [np.arange(N) if p is None else np.asarray(p) for p in perm]
    return lens, shifts, full


PINS = {
    "crossing_T": "closed forms + unitarity/|det| invariants + ideal cross/bar special cases (tests/test_oracle.py)",
    "crossing_dT": "central finite differences of crossing_T in FP64 (tests/test_oracle.py)",
    "grad_from_dJdT": "finite differences of a real loss through the whole mesh (tests/test_oracle.py)",
    "fwd": "golden vectors from the compiled reference c/dot.cpp mode 0 + dense brute force",
    "jvp": "golden vectors from c/dot.cpp mode 1 + central finite differences",
    "vjp_inverse": "golden vectors from c/dot.cpp mode 2 (threads=1) + adjoint identity + FD",
    "vjp_stored": "equals vjp_inverse on unitary meshes + FD on lossy meshes",
    "Geometry(fft)": "dense block-diagonal construction of gpu/test.py:73-77 restated + FD",
    "structured_run": "golden vectors from c/dot.cpp modes 0/1/2 with permutations and the phase screen at "
                      "either position + dense brute force (structured_dense) + FD of the screen phases",
    "structured_dense": "agrees with structured_run mode 0 and with the reference engine's golden vectors",
    "butterfly_perms": "permutation pairs checked against the stride-pairing definition (adjacent after gather)",
    "crossing_d2T": "central finite differences of crossing_dT (all four types)",
    "hvp_inverse": "central finite differences of the FP64 gradient (mesh_vjp) along q; its gradient equals mesh_vjp",
}
