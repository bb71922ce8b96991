"""Exact Hessian-vector products (forward-over-reverse kernel) against the FP64 oracle.

The oracle's hvp_inverse is pinned by central differences of its gradient; here the GPU result
(meshprop_hvp after fwddiff along q) must match it within 1e-5 relative L2: MZI and SYM on NN
meshes up to N = 1024 (every NN tile width, W = 2 .. 16) and on FFT meshes up to N = 512, orth
(real fields) on NN meshes, several vector tiles per workgroup; MeshNetworkGPU.hvp_target /
hess_target must agree with the oracle and be symmetric; meshes with a phase screen (input or output)
against FP64 differences of the oracle's gradient.
"""
import zlib

import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
CASES = [("mzi", "nn", 16, 9, 7), ("mzi", "clements", 64, 0, 20), ("sym", "nn", 40, 11, 13),
         ("mzi", "fft", 128, 7, 10), ("mzi", "clements", 256, 0, 6), ("mzi", "nn", 512, 12, 5),
         ("sym", "fft", 256, 8, 9), ("mzi", "nn", 130, 5, 4200), ("sym", "nn", 300, 10, 70),
         ("mzi", "nn", 700, 9, 6), ("sym", "nn", 1000, 7, 5), ("mzi", "clements", 1024, 0, 3),
         ("orth", "nn", 96, 14, 9), ("orth", "clements", 128, 0, 8), ("orth", "nn", 250, 11, 33),
         ("orth", "nn", 600, 8, 4), ("orth", "nn", 1024, 6, 5), ("mzi", "clements", 2, 0, 3),
         ("sym", "clements", 6, 0, 1), ("mzi", "nn", 514, 1, 9), ("orth", "clements", 4, 0, 5),
         ("mzi", "clements", 130, 0, 2049)]


def _geom(rng, kind, N, L):
    if kind == "clements":
        return O.clements(N)
    if kind == "fft":
        return O.Geometry(N, L, strides=rng.randint(0, int(np.log2(N)), size=L))
    return O.random_geometry_nn(rng, N, L)


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c)) for c in CASES])
def test_hvp_kernel_matches_oracle(gpu, case):
    import meshes_amd._lib as Lb
    from meshes_amd import MeshNetworkGPU
    X, kind, N, L, B = case
    rng = np.random.RandomState(zlib.crc32(str(case).encode()) % 10000)
    g = _geom(rng, kind, N, L)
    n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
    s = (0.05 * rng.standard_normal((g.L, N // 2, n_s))).astype(np.float32)
    q = rng.standard_normal(p.shape).astype(np.float32)
    real = O.REAL_FIELD[X]
    fdt = np.float32 if real else np.complex64
    u = O.unit_cn(rng, B, N, real=real).astype(fdt)
    t = O.unit_cn(rng, B, N, real=real).astype(fdt)
    kw = dict(strides=g.strides) if g.kind == "fft" else {}
    mesh = MeshNetworkGPU(N, g.L, g.lens, g.shifts, p, s, X, is_phase=False, **kw)
    assert mesh.exact_hvp_supported()
    out, dout = mesh.dot(torch.as_tensor(u.T), dp=q)
    qt = torch.as_tensor(q.reshape(-1)).cuda()
    grad = torch.zeros(mesh.n_phase, dtype=torch.float32, device="cuda")
    hv = torch.zeros_like(grad)
    gin = torch.full((B, N), float("nan"), dtype=mesh.dtype, device="cuda")
    dgin = torch.full_like(gin, float("nan"))
    nb = Lb.hvp_workspace_size(X, N, g.L, B, mesh.fft)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    g0, g1 = mesh._geom()
    Lb.hvp(X, N, g.L, B, mesh.fft, g0, g1, mesh.p_phase, qt, N // 2 * n_p, mesh.p_splitter, N // 2 * n_s, out.T,
           dout.T, None, None, torch.as_tensor(t).cuda(), grad, hv, N, ws, nb, dJdu_in=gin, d_dJdu_in=dgin)
    rg, rh = O.hvp_inverse(g, X, p.astype(np.float64), s.astype(np.float64), u.astype(np.complex128),
                           q.astype(np.float64), t.astype(np.complex128))
    assert O.rel_l2(grad.cpu().numpy().reshape(rg.shape), rg) < TOL
    assert O.rel_l2(hv.cpu().numpy().reshape(rh.shape), rh) < TOL
    # the input gradient (oracle adjoint) and its tangent along q (FP64 central difference)
    s64, u64, t64 = s.astype(np.float64), u.astype(np.complex128), t.astype(np.complex128)

    def g_input(pp):
        o = O.mesh_fwd(g, X, pp, s64, u64)
        return O.mesh_vjp(g, X, pp, s64, o, 2 * (o - t64), u_in=u64)[1]
    h = 1e-5
    p64, q64 = p.astype(np.float64), q.astype(np.float64)
    assert O.rel_l2(gin.cpu().numpy(), g_input(p64)) < TOL
    fd = (g_input(p64 + h * q64) - g_input(p64 - h * q64)) / (2 * h)
    assert O.rel_l2(dgin.cpu().numpy(), fd) < TOL


def test_hvp_target_and_hessian(gpu):
    """MeshNetworkGPU.hvp_target (exact path) against the oracle's forward-over-reverse of the
    L2norm objective, and hess_target symmetric."""
    from meshes_amd import MeshNetworkGPU
    rng = np.random.RandomState(5)
    N = 8
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    mesh = MeshNetworkGPU(N, N, g.lens, g.shifts, p, np.zeros(0), "mzi", is_phase=False)
    assert mesh.exact_hvp_supported()
    U = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))[0].astype(np.complex64)
    q = rng.standard_normal(mesh.n_phase).astype(np.float32)
    hv = mesh.hvp_target(q, U).cpu().numpy()
    # FP64 reference: J(p) = |offdiag(U_mesh U^H)|^2, H q by central differences of its gradient
    v = U.astype(complex).conj()          # rows = columns of U^H

    def grad(pp):
        out = O.mesh_fwd(g, "mzi", pp, None, v)
        V = out - np.eye(N)
        V = V - np.diag(np.diag(V))
        return O.mesh_vjp(g, "mzi", pp, None, out, 2 * V)[0].reshape(-1)
    h = 1e-5
    p64, q64 = p.astype(np.float64), q.astype(np.float64).reshape(p.shape)
    fd = (grad(p64 + h * q64) - grad(p64 - h * q64)) / (2 * h)
    assert O.rel_l2(hv, fd) < 1e-4
    H = mesh.hess_target(U).cpu().numpy()
    assert O.rel_l2(H, H.T) < 1e-4


def _quartic_case():
    from meshes_amd import MeshNetworkGPU
    rng = np.random.RandomState(9)
    N, B = 16, 12
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    mesh = MeshNetworkGPU(N, N, g.lens, g.shifts, p, np.zeros(0), "mzi", is_phase=False)
    v = O.unit_cn(rng, B, N).astype(np.complex64)
    q = rng.standard_normal(mesh.n_phase).astype(np.float32)

    def grad(pp):  # FP64 gradient of J = sum |y|^4 (output gradient 4 |y|^2 y)
        out = O.mesh_fwd(g, "mzi", pp, None, v.astype(np.complex128))
        return O.mesh_vjp(g, "mzi", pp, None, out, 4 * np.abs(out) ** 2 * out)[0].reshape(-1)
    h = 1e-5
    p64, q64 = p.astype(np.float64), q.astype(np.float64).reshape(p.shape)
    fd = (grad(p64 + h * q64) - grad(p64 - h * q64)) / (2 * h)
    return mesh, v, q, fd


def test_hvp_non_affine_output_gradient(gpu):
    """hvp with a nonlinear (but traceable) w: w(y) = 4 |y|^2 y, the output gradient of
    J = sum |y|^4.  The exact path takes w's tangent from torch.func.jvp; checked against an FP64
    central difference of the oracle's gradient."""
    mesh, v, q, fd = _quartic_case()
    assert mesh.exact_hvp_supported()

    def w(y):
        return 4 * (y * y.conj()).real * y
    hv = mesh.hvp(q, torch.as_tensor(v.T).cuda(), w).cpu().numpy()
    assert O.rel_l2(hv, fd) < TOL


def test_hvp_untraceable_w_falls_back_to_gradient_difference(gpu):
    """A w that leaves torch (through NumPy) cannot be traced by torch.func.jvp: hvp then takes
    the reference's central difference of grad_phi (gpu/mesh.py:236-251), accurate for any w
    (FP32 gradients, eps = 1e-3: the reference's own 1e-2 gradient bar, gpu/test.py:154)."""
    mesh, v, q, fd = _quartic_case()

    def w(y):
        a = y.detach().cpu().numpy()
        return torch.as_tensor(4 * np.abs(a) ** 2 * a, device=y.device)
    hv = mesh.hvp(q, torch.as_tensor(v.T).cuda(), w).cpu().numpy()
    assert O.rel_l2(hv, fd) < 1e-2


def _screen_grad(g, X, p, ph, v, gfun, pos):
    """FP64 gradient (mesh phases, then screen phases) of a loss with dJ/dout = gfun(out), the
    screen exp(i ph) at the mesh input or output; the oracle's mesh adjoint composed with the
    screen's, whose own gradient is pinned against loss differences with the screen tests."""
    e = np.exp(1j * ph)
    if pos == "in":
        xs = v * e[None, :]
        out = O.mesh_fwd(g, X, p, None, xs)
        rdp, gxs, _ = O.mesh_vjp(g, X, p, None, out, gfun(out), u_in=xs)
        dph = np.sum(np.imag(gxs * np.conj(xs)), axis=0)
    else:
        z = O.mesh_fwd(g, X, p, None, v)
        out = z * e[None, :]
        gout = gfun(out)
        dph = np.sum(np.imag(gout * np.conj(out)), axis=0)
        rdp = O.mesh_vjp(g, X, p, None, z, gout * np.conj(e)[None, :], u_in=v)[0]
    return np.concatenate([rdp.reshape(-1), dph])


SCREEN_CASES = [("mzi", "clements", 16, 0, 12, "out"), ("sym", "nn", 64, 9, 11, "out"),
                ("mzi", "fft", 32, 5, 7, "out"), ("mzi", "nn", 40, 6, 5, "in"), ("sym", "clements", 16, 0, 9, "in"),
                ("mzi", "fft", 64, 6, 4, "in")]


@pytest.mark.parametrize("case", SCREEN_CASES, ids=["-".join(map(str, c)) for c in SCREEN_CASES])
def test_hvp_with_phase_screen(gpu, case):
    """hvp and hvp_target on meshes with a phase screen, against FP64 central differences of the
    oracle's gradient (screen phases included), both screen positions on the exact path (1e-5)."""
    from meshes_amd import MeshNetworkGPU
    X, kind, N, L, B, pos = case
    rng = np.random.RandomState(zlib.crc32(repr(case).encode()))
    g = _geom(rng, kind, N, L)
    n_p = O.N_PHASE[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
    ph = rng.uniform(0, 2 * np.pi, N).astype(np.float32)
    mesh = MeshNetworkGPU(N, g.L, g.lens, g.shifts, np.concatenate([p.ravel(), ph]), np.zeros(0), X, phi_pos=pos,
                          is_phase=True, strides=g.strides if kind == "fft" else None)
    assert mesh.exact_hvp_supported()
    tol = TOL
    v = O.unit_cn(rng, B, N).astype(np.complex64)
    q = rng.standard_normal(mesh.n_phase).astype(np.float32)
    h = 1e-5
    p64, ph64 = p.astype(np.float64), ph.astype(np.float64)
    qm, qs = q[:p.size].astype(np.float64).reshape(p.shape), q[p.size:].astype(np.float64)

    def fd(gfun, vin):
        return (_screen_grad(g, X, p64 + h * qm, ph64 + h * qs, vin, gfun, pos)
                - _screen_grad(g, X, p64 - h * qm, ph64 - h * qs, vin, gfun, pos)) / (2 * h)

    def w(y):  # J = sum |y|^4
        return 4 * (y * y.conj()).real * y
    hv = mesh.hvp(q, torch.as_tensor(v.T).cuda(), w).cpu().numpy()
    assert O.rel_l2(hv, fd(lambda o: 4 * np.abs(o) ** 2 * o, v.astype(np.complex128))) < tol
    # L2norm keeps the diagonal with a screen: dJ/dU = 2 (U - 1)
    U = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))[0].astype(np.complex64)
    hv = mesh.hvp_target(q, U).cpu().numpy()
    assert O.rel_l2(hv, fd(lambda o: 2 * (o - np.eye(N)), U.astype(np.complex128).conj())) < tol
