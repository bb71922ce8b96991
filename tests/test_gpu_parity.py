"""GPU parity: libmeshprop (HIP, gfx950) against the FP64 oracle, through the C ABI.

Bar: complex64 / float32 outputs and gradients within 1e-5 relative L2 of the FP64
oracle evaluated on the same FP32-rounded inputs and parameters.
Geometries and edge cases follow gpu/test.py:24-157 (random NN shifts/lens, N below the
padded width, ragged batches, random tiles) plus FFT strides, all four crossing models,
in-place backdiff (gpu/mesh.py:228-231) and a bitwise-repeatability check.
"""
import zlib

import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _setup(X, kind, N, L, B, seed, errors=True, lossy=False, clem=False):
    rng = np.random.RandomState(seed)
    if kind == "fft":
        n = int(np.log2(N))
        g = O.Geometry(N, L, strides=rng.randint(0, n, size=L))
    elif clem:
        g = O.clements(N)
        L = N
    else:
        g = O.random_geometry_nn(rng, N, L)
    n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
    s = None
    if n_s and errors:
        s = (0.05 * rng.standard_normal((g.L, N // 2, n_s))).astype(np.float32)
        if X == "mzi_generic":
            s[..., 10:] = np.float32(np.log(10 ** (-0.1 / 20)) * 2) if lossy else 0.0
    real = O.REAL_FIELD[X]
    u = O.unit_cn(rng, B, N, real=real)
    u = u.astype(np.float32 if real else np.complex64)
    return g, p, s, u, rng


def _mesh(g, X, p, s, tile=None):
    from meshes_amd import MeshNetworkGPU
    kw = dict(strides=g.strides) if g.kind == "fft" else {}
    return MeshNetworkGPU(g.N, g.L, g.lens, g.shifts, p, s if s is not None else np.zeros(0), X,
                          is_phase=False, tile=tile, **kw)


def _np(t):
    return t.detach().cpu().numpy()


def _f64(s):
    return None if s is None else s.astype(np.float64)


CASES = [
    ("mzi", "nn", 256, 12, 37), ("mzi", "nn", 250, 9, 5), ("mzi", "nn", 64, 20, 70), ("mzi", "nn", 16, 16, 64),
    ("mzi", "nn", 2, 3, 9), ("mzi", "nn", 6, 5, 11), ("mzi", "nn", 130, 7, 33), ("mzi", "nn", 512, 6, 20),
    ("mzi", "nn", 1000, 4, 9), ("mzi", "nn", 1024, 3, 5), ("mzi", "fft", 256, 10, 21), ("mzi", "fft", 32, 8, 40),
    ("mzi", "fft", 1024, 10, 6), ("sym", "nn", 200, 8, 17), ("sym", "fft", 128, 7, 19), ("orth", "nn", 256, 10, 41),
    ("orth", "nn", 38, 9, 13), ("orth", "fft", 512, 9, 8), ("mzi_generic", "nn", 96, 8, 12),
    ("orth", "nn", 700, 11, 19), ("sym", "nn", 600, 21, 70),
]


def _id(c):
    return "-".join(map(str, c))


def _seed(c):
    return zlib.crc32(_id(c).encode()) % 10000


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_fwdprop(gpu, case):
    X, kind, N, L, B = case
    g, p, s, u, _ = _setup(X, kind, N, L, B, seed=_seed(case))
    mesh = _mesh(g, X, p, s)
    out = _np(mesh.dot(torch.as_tensor(u.T)))
    ref = O.mesh_fwd(g, X, p.astype(np.float64), _f64(s), u.astype(np.complex128))
    assert O.rel_l2(out.T, ref) < TOL


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_fwddiff(gpu, case):
    X, kind, N, L, B = case
    g, p, s, u, rng = _setup(X, kind, N, L, B, seed=1 + _seed(case))
    dp = rng.standard_normal(p.shape).astype(np.float32)
    du = (rng.standard_normal(u.shape) + (0 if O.REAL_FIELD[X] else 1j * rng.standard_normal(u.shape)))
    du = du.astype(u.dtype)
    mesh = _mesh(g, X, p, s)
    out, dout = mesh.dot(torch.as_tensor(u.T), dp=dp, dv=torch.as_tensor(du.T))
    ref, dref = O.mesh_jvp(g, X, p.astype(np.float64), _f64(s), u.astype(np.complex128), dp.astype(np.float64),
                           du.astype(np.complex128))
    assert O.rel_l2(_np(out).T, ref) < TOL
    assert O.rel_l2(_np(dout).T, dref) < TOL


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_backdiff(gpu, case):
    X, kind, N, L, B = case
    g, p, s, u, rng = _setup(X, kind, N, L, B, seed=2 + _seed(case))
    mesh = _mesh(g, X, p, s)
    uo = _np(mesh.dot(torch.as_tensor(u.T))).T
    y = O.unit_cn(rng, B, N, real=O.REAL_FIELD[X]).astype(u.dtype)
    gout = (2 * (uo - y)).astype(u.dtype)
    dp, gin = mesh.grad_phi(torch.as_tensor(uo.T), torch.as_tensor(gout.T), vpos="out")
    lossless = X != "mzi_generic"
    rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), uo.astype(np.complex128),
                              gout.astype(np.complex128), u_in=None if lossless else u.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin).T, rgin) < TOL


def test_backdiff_reconstructs_input_in_place(gpu):
    """backdiff with u_in aliasing u_out (as gpu/mesh.py:228-231 calls it) recovers the input."""
    import meshes_amd._lib as Lb
    g, p, s, u, rng = _setup("mzi", "nn", 256, 16, 29, seed=5)
    mesh = _mesh(g, "mzi", p, s)
    ub = torch.as_tensor(u).cuda().contiguous()
    out = torch.empty_like(ub)
    ws, nb = mesh._workspace(Lb.OP_FWDPROP, 29)
    Lb.run(Lb.OP_FWDPROP, "mzi", 256, g.L, 29, False, mesh.lens, mesh.shifts, mesh.p_phase, None, 256,
           mesh.p_splitter, 256, ub, None, out, None, 256, workspace=ws, ws_bytes=nb)
    torch.cuda.synchronize()
    y = torch.as_tensor(O.unit_cn(rng, 29, 256).astype(np.complex64)).cuda()
    gb = (2 * (out - y)).contiguous()
    gref = 2 * (_np(out).astype(np.complex128) - _np(y))
    dp = torch.zeros_like(mesh.p_phase)
    ws, nb = mesh._workspace(Lb.OP_BACKDIFF, 29)
    Lb.run(Lb.OP_BACKDIFF, "mzi", 256, g.L, 29, False, mesh.lens, mesh.shifts, mesh.p_phase, dp, 256,
           mesh.p_splitter, 256, out, gb, out, gb, 256, workspace=ws, ws_bytes=nb)
    torch.cuda.synchronize()
    assert O.rel_l2(_np(out), u) < TOL
    rdp, rgin, _ = O.mesh_vjp(g, "mzi", p.astype(np.float64), _f64(s),
                              O.mesh_fwd(g, "mzi", p.astype(np.float64), _f64(s), u.astype(np.complex128)), gref)
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gb), rgin) < TOL


@pytest.mark.parametrize("tile", [(2, 1, 4), (4, 1, 1), (4, 2, 2), (4, 4, 4), (8, 1, 4), (8, 2, 3), (16, 1, 4),
                                  (4, 8, 4), (8, 4, 2)])
def test_tiles(gpu, tile):
    """Every compiled tile gives the same answer (N=200 < padded width, ragged batch)."""
    g, p, s, u, rng = _setup("mzi", "nn", 200, 9, 45, seed=17)
    mesh = _mesh(g, "mzi", p, s, tile=tile)
    out = _np(mesh.dot(torch.as_tensor(u.T))).T
    ref = O.mesh_fwd(g, "mzi", p.astype(np.float64), _f64(s), u.astype(np.complex128))
    assert O.rel_l2(out, ref) < TOL
    gout = (out - u).astype(np.complex64)
    dp, gin = mesh.grad_phi(torch.as_tensor(out.T), torch.as_tensor(gout.T), vpos="out")
    rdp, rgin, _ = O.mesh_vjp(g, "mzi", p.astype(np.float64), _f64(s), out.astype(np.complex128),
                              gout.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin).T, rgin) < TOL


def test_golden_c1_through_gpu(gpu):
    """Config C1 inputs (the reference engine's golden vectors) through the GPU path."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_clements16.npz"))
    g = O.Geometry(16, 16, lens=d["lens"], shifts=d["shifts"])
    p = d["p"].astype(np.float32)
    mesh = _mesh(g, "mzi", p, None)
    u = d["u_in"].astype(np.complex64)
    out = _np(mesh.dot(torch.as_tensor(u.T))).T
    assert O.rel_l2(out, d["fwd_out"]) < TOL  # fixture computed from the FP64 parameters
    gout = d["g_out"].astype(np.complex64)
    dp, gin = mesh.grad_phi(torch.as_tensor(d["fwd_out"].astype(np.complex64).T), torch.as_tensor(gout.T),
                            vpos="out")
    assert O.rel_l2(_np(gin).T, d["vjp_g_in"]) < TOL
    rdp = O.grad_from_dJdT("mzi", d["p"], None, d["vjp_dJdT"])
    rdp[~O.active_mask(g)] = 0
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < 2 * TOL


def test_deterministic_gradient(gpu):
    g, p, s, u, rng = _setup("mzi", "nn", 256, 24, 700, seed=23)
    mesh = _mesh(g, "mzi", p, s)
    ut = torch.as_tensor(u.T)
    gt = torch.as_tensor((0.5 * u).T)
    a = mesh.grad_phi(ut, gt)[0].clone()
    b = mesh.grad_phi(ut, gt)[0].clone()
    assert torch.equal(a, b)


def test_edge_sizes(gpu):
    """L = 0 copies; B = 1; lens = 0 layers are the identity."""
    from meshes_amd import MeshNetworkGPU
    rng = np.random.RandomState(3)
    u = O.unit_cn(rng, 3, 8).astype(np.complex64)
    m0 = MeshNetworkGPU(8, 0, [], [], np.zeros(0), np.zeros(0), "mzi", is_phase=False)
    assert np.allclose(_np(m0.dot(torch.as_tensor(u.T))).T, u)
    m1 = MeshNetworkGPU(8, 2, [0, 0], [0, 1], rng.standard_normal(16), np.zeros(0), "mzi", is_phase=False)
    assert np.allclose(_np(m1.dot(torch.as_tensor(u.T))).T, u)
    g, p, s, u1, _ = _setup("mzi", "nn", 64, 5, 1, seed=4)
    mesh = _mesh(g, "mzi", p, s)
    ref = O.mesh_fwd(g, "mzi", p.astype(np.float64), _f64(s), u1.astype(np.complex128))
    assert O.rel_l2(_np(mesh.dot(torch.as_tensor(u1.T))).T, ref) < TOL


def test_c2_full_size(gpu):
    """Config C2 (Clements MZI, N=256, L=256, B=4096, ideal splitters): oracle parity on a
    256-vector slice (forward, JVP and gradient) and size-independent checks on the whole
    batch: unitarity, exact reconstruction of the input, the fused L2 gradient equal to the
    explicit one, and the adjoint identity <g, J dp> = <dp_grad, dp> between the fwddiff and backdiff kernels."""
    N = 256
    rng = np.random.RandomState(2)
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    u = O.unit_cn(np.random.RandomState(4), 4096, N).astype(np.complex64)
    pt = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    mesh, target = _mesh(g, "mzi", p, None), _mesh(g, "mzi", pt, None)
    ut = torch.as_tensor(u.T).cuda()
    y = target.dot(ut)
    out = mesh.dot(ut)
    # unitarity per vector
    nrm = torch.linalg.norm(out, dim=0)
    assert torch.max(torch.abs(nrm - 1)).item() < 1e-5
    # oracle on a slice
    sl = slice(0, 256)
    ref = O.mesh_fwd(g, "mzi", p.astype(np.float64), None, u[sl].astype(np.complex128))
    assert O.rel_l2(_np(out).T[sl], ref) < TOL
    # fused gradient == explicit gradient; reconstruction
    J, dp_f, uo = mesh.grad_target_l2(ut, y)
    dp_e, gin = mesh.grad_phi(out, 2 * (out - y), vpos="out")
    assert O.rel_l2(_np(dp_f), _np(dp_e)) < 1e-6
    # slice gradient vs oracle
    dps, _ = mesh.grad_phi(out[:, sl], 2 * (out - y)[:, sl], vpos="out")
    rdp, _, _ = O.mesh_vjp(g, "mzi", p.astype(np.float64), None, _np(out[:, sl]).T.astype(np.complex128),
                           2 * (_np(out[:, sl]) - _np(y[:, sl])).T.astype(np.complex128))
    assert O.rel_l2(_np(dps).reshape(rdp.shape), rdp) < TOL
    # adjoint identity along a random unit direction over the active slots
    d = rng.standard_normal(p.shape)
    d[~O.active_mask(g)] = 0
    d = (d / np.linalg.norm(d)).astype(np.float32)
    _, dout = mesh.dot(ut, dp=d)
    # the JVP (fwddiff) itself against the FP64 oracle on the slice
    _, dref = O.mesh_jvp(g, "mzi", p.astype(np.float64), None, u[sl].astype(np.complex128), d.astype(np.float64))
    assert O.rel_l2(_np(dout).T[sl], dref) < TOL
    lhs = torch.sum(torch.real(torch.conj(2 * (out - y)) * dout)).item()
    rhs = float(np.sum(_np(dp_e).astype(np.float64) * d.reshape(-1)))
    assert abs(lhs - rhs) <= 1e-4 * max(abs(rhs), 1e-3)


def test_unitary_checkpointed_adjoint_w8(gpu):
    """Unitary MZI at N=512 (W=8 pipelined K-form stage, C=2): backdiff with checkpoints every 8
    layers in one launch and every 5 layers (one launch per segment) equals the oracle."""
    import meshes_amd._lib as Lb
    X, N, L, B = "mzi", 512, 30, 300
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=77)
    mesh = _mesh(g, X, p, s)
    ldp, lds = mesh._ld()
    ub = torch.as_tensor(u).cuda().contiguous()
    y = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    for K in (8, 5):
        out = torch.empty_like(ub)
        ck = torch.empty(((L + K - 1) // K, B, N), dtype=ub.dtype, device=ub.device)
        ws, nb = mesh._workspace(Lb.OP_FWDPROP, B)
        Lb.run(Lb.OP_FWDPROP, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, None, ldp, mesh.p_splitter,
               lds, ub, None, out, None, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
        gb = (2 * (out - y)).contiguous()
        dp, uin, gin = torch.zeros_like(mesh.p_phase), torch.empty_like(ub), torch.empty_like(ub)
        ws, nb = mesh._workspace(Lb.OP_BACKDIFF, B)
        Lb.run(Lb.OP_BACKDIFF, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, dp, ldp, mesh.p_splitter,
               lds, out, gb, uin, gin, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
        torch.cuda.synchronize()
        rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128),
                                  _np(gb).astype(np.complex128), u_in=u.astype(np.complex128))
        assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
        assert O.rel_l2(_np(gin), rgin) < TOL
        assert O.rel_l2(_np(uin), u) < TOL


@pytest.mark.parametrize("N,K", [(64, 8), (256, 8), (256, 5), (512, 8)])
def test_lossy_checkpointed_adjoint(gpu, N, K):
    """Lossy generic MZI (non-unitary): backdiff with forward checkpoints every K layers equals
    the oracle's stored-field adjoint.  N=256 runs the pipelined adjoint stage (C=2): K=8 in one
    launch, K=5 (not a multiple of the stage) one launch per segment; N=512 with this batch runs
    the one-vector W=8 tile in its lean form."""
    import meshes_amd._lib as Lb
    X, L, B = "mzi_generic", 40, 50
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=31, lossy=True)
    g = O.clements(N)
    L = N
    p = rng.uniform(0, 2 * np.pi, (L, N // 2, 2)).astype(np.float32)
    s = np.zeros((L, N // 2, 14), np.float32)
    s[..., :2] = 0.02 * rng.standard_normal((L, N // 2, 2))
    s[..., 10:] = np.float32(np.log(10 ** (-0.1 / 20)))
    mesh = _mesh(g, X, p, s)
    ub = torch.as_tensor(u).cuda().contiguous()
    out = torch.empty_like(ub)
    nck = (L + K - 1) // K
    ck = torch.empty((nck, B, N), dtype=ub.dtype, device=ub.device)
    ws, nb = mesh._workspace(Lb.OP_FWDPROP, B)
    Lb.run(Lb.OP_FWDPROP, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, None, N, mesh.p_splitter, N * 7,
           ub, None, out, None, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
    y = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    gb = (2 * (out - y)).contiguous()
    dp = torch.zeros_like(mesh.p_phase)
    uin = torch.empty_like(ub)
    gin = torch.empty_like(ub)
    ws, nb = mesh._workspace(Lb.OP_BACKDIFF, B)
    Lb.run(Lb.OP_BACKDIFF, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, dp, N, mesh.p_splitter, N * 7,
           out, gb, uin, gin, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
    torch.cuda.synchronize()
    rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), s.astype(np.float64), _np(out).astype(np.complex128),
                              _np(gb).astype(np.complex128), u_in=u.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL
    assert O.rel_l2(_np(uin), u) < TOL
