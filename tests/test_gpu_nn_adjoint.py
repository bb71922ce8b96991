"""GPU parity of the adjoint kernel specialised for unitary crossings on NN geometries.

backdiff and fwd+grad of MZI and symmetric crossings (and scaled-unitary generic MZIs) on NN
meshes run k_bwd_nn: the pipelined K-form adjoint with the workgroup's wave count and stage
depth compiled in.  It performs the same arithmetic in the same fixed summation orders as the
general adjoint (MESHPROP_GENERIC_ADJOINT selects that one), so every output must be bitwise
equal to the general kernel's when both run the same workgroup shape, and within 1e-5 relative
L2 of the FP64 oracle (gpu/mesh.py:208-234 semantics).  The widest tile (W = 16, N > 512) runs
one-layer stages in stage order with four- or eight-wave workgroups (the general kernel has four:
eight-wave cases are checked against the oracle).  Cases cover each compiled width (W = 2, 4, 8,
16: N up to 128, 256, 512, 1024), one,
two and four vectors per lane, four- and eight-wave workgroups, depths that are not a multiple
of the stage (a partial last stage, odd or even), several vector tiles per workgroup, random
shifts and lengths (gpu/test.py:41), padded widths (N below 64 W) and checkpointed
scaled-unitary adjoints.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O
from tests.test_gpu_parity import _f64, _mesh, _np, _setup

pytestmark = pytest.mark.gpu

TOL = 1e-5

# (X, N, L, B, tile): tiles (w, v, waves) force each compiled variant; None = the library's choice.
CASES = [
    ("mzi", 256, 256, 600, None), ("mzi", 256, 37, 300, (4, 2, 8)), ("mzi", 250, 30, 129, (4, 1, 4)),
    ("mzi", 200, 13, 70, (4, 4, 4)), ("mzi", 128, 41, 500, (2, 1, 8)), ("mzi", 72, 9, 45, (2, 2, 4)),
    ("mzi", 100, 19, 90, (2, 4, 8)), ("mzi", 512, 22, 200, None), ("mzi", 300, 7, 33, (8, 1, 4)),
    ("mzi", 400, 11, 64, (8, 2, 8)), ("sym", 256, 21, 150, None), ("sym", 128, 18, 77, (2, 2, 8)),
    ("mzi", 130, 3, 5, None), ("mzi", 256, 1, 40, (4, 2, 8)),
]
# W = 16: the library's choice, forced four- and eight-wave workgroups, padded widths, odd depths.
WIDE = [
    ("mzi", 1024, 37, 100, None), ("mzi", 1024, 20, 64, (16, 1, 4)), ("mzi", 1000, 13, 70, (16, 1, 8)),
    ("mzi", 600, 9, 33, (16, 1, 4)), ("sym", 1024, 11, 64, (16, 1, 8)), ("sym", 770, 1, 9, (16, 1, 4)),
    ("mzi", 1024, 1024, 40, (16, 1, 8)), ("mzi", 1024, 16, 2048, None),
]


def _id(c):
    return "-".join(str(x) for x in c[:4]) + ("" if c[4] is None else "-t" + "x".join(map(str, c[4])))


def _call(mesh, op, B, u_a, du_a, u_b, du_b, dp, target, tile, flags, ckpt=None, K=0):
    import meshes_amd._lib as Lb
    ldp, lds = mesh._ld()
    g0, g1 = mesh._geom()
    s = mesh.p_splitter if mesh.p_splitter.numel() else None
    nb = Lb.workspace_size(op, mesh.X, mesh.N, mesh.L, B, False, tile)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=u_a.device)
    Lb.run(op, mesh.X, mesh.N, mesh.L, B, False, g0, g1, mesh.p_phase, dp, ldp, s, lds, u_a, du_a, u_b, du_b, mesh.N,
           target=target, ckpt=ckpt, ckpt_every=K, tile=tile, workspace=ws, ws_bytes=nb, flags=flags)


def _same_shape(X, N, L, B, tile):
    """Whether the specialised and the general kernel run the same workgroup shape (the
    general W = 16 adjoint has four-wave workgroups only)."""
    import meshes_amd._lib as Lb
    w, _, waves, _ = Lb.resolve_tile(Lb.OP_BACKDIFF, X, N, B, tile)
    return not (w == 16 and waves > 4)


@pytest.mark.parametrize("case", CASES + WIDE, ids=[_id(c) for c in CASES + WIDE])
def test_nn_adjoint_backdiff(gpu, case):
    """backdiff (reconstructed input, input adjoint, dp accumulated) against the general
    kernel bitwise (same workgroup shape) and the oracle."""
    import meshes_amd._lib as Lb
    X, N, L, B, tile = case
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=7 * N + L)
    mesh = _mesh(g, X, p, s)
    tile = None if tile is None else tuple(tile) + (0,)
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, X, N, L, B, tile=tile) == "nn"
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, X, N, L, B, tile=tile, flags=Lb.GENERIC_ADJOINT) == "general"
    ub = torch.as_tensor(u).cuda().contiguous()
    out = mesh.dot(ub.T).T.contiguous()
    gout = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().contiguous()
    res = {}
    for flags in (0, Lb.GENERIC_ADJOINT):
        dp = torch.full_like(mesh.p_phase, 0.5)  # accumulates into what the caller holds
        uin, gin = torch.empty_like(ub), torch.empty_like(ub)
        _call(mesh, Lb.OP_BACKDIFF, B, out, gout, uin, gin, dp, None, tile, flags)
        res[flags] = (dp, uin, gin)
    torch.cuda.synchronize()
    if _same_shape(X, N, L, B, tile):
        for a, b in zip(res[0], res[Lb.GENERIC_ADJOINT]):
            assert torch.equal(a, b)
    dp, uin, gin = res[0]
    rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128),
                              _np(gout).astype(np.complex128))
    assert O.rel_l2(_np(dp - 0.5).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL
    assert O.rel_l2(_np(uin), u) < TOL


@pytest.mark.parametrize("case", CASES[:8] + WIDE, ids=[_id(c) for c in CASES[:8] + WIDE])
def test_nn_adjoint_fwdgrad(gpu, case):
    """One-call fwd+grad with the fused L2 target: bitwise equal to the general kernel."""
    import meshes_amd._lib as Lb
    X, N, L, B, tile = case
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=3 * N + L)
    mesh = _mesh(g, X, p, s)
    tile = None if tile is None else tuple(tile) + (0,)
    assert Lb.adjoint_kernel(Lb.OP_FWDGRAD, X, N, L, B, tile=tile) == "nn"
    ub = torch.as_tensor(u).cuda().contiguous()
    yb = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().contiguous()
    res = {}
    for flags in (0, Lb.GENERIC_ADJOINT):
        dp = torch.zeros_like(mesh.p_phase)
        out, gin = torch.empty_like(ub), torch.empty_like(ub)
        _call(mesh, Lb.OP_FWDGRAD, B, ub, None, out, gin, dp, yb, tile, flags | Lb.DP_ASSIGN)
        res[flags] = (dp, out, gin)
    torch.cuda.synchronize()
    if _same_shape(X, N, L, B, tile):
        for a, b in zip(res[0], res[Lb.GENERIC_ADJOINT]):
            assert torch.equal(a, b)
    else:
        assert torch.equal(res[0][1], res[Lb.GENERIC_ADJOINT][1])  # the forward is the same kernel
    dp, out, _ = res[0]
    gout = 2 * (_np(out).astype(np.complex128) - _np(yb))
    rdp, _, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128), gout)
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL


@pytest.mark.parametrize("X,N,L,B,tile", [("mzi", 128, 10, 2300, (2, 1, 4, 0)), ("mzi", 1024, 4, 2500, (16, 1, 4, 0))])
def test_nn_adjoint_several_tiles_per_workgroup(gpu, X, N, L, B, tile):
    """More vector tiles than workgroups (the grid is capped at two workgroups per CU): each
    workgroup adds its later tiles into its partial."""
    import meshes_amd._lib as Lb
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=5)
    mesh = _mesh(g, X, p, s)
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, X, N, L, B, tile=tile) == "nn"
    ub = torch.as_tensor(u).cuda().contiguous()
    out = mesh.dot(ub.T).T.contiguous()
    gout = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().contiguous()
    res = {}
    for flags in (0, Lb.GENERIC_ADJOINT):
        dp = torch.zeros_like(mesh.p_phase)
        uin, gin = torch.empty_like(ub), torch.empty_like(ub)
        _call(mesh, Lb.OP_BACKDIFF, B, out, gout, uin, gin, dp, None, tile, flags)
        res[flags] = (dp, gin)
    torch.cuda.synchronize()
    assert torch.equal(res[0][0], res[Lb.GENERIC_ADJOINT][0]) and torch.equal(res[0][1], res[Lb.GENERIC_ADJOINT][1])
    rdp, _, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128),
                           _np(gout).astype(np.complex128))
    assert O.rel_l2(_np(res[0][0]).reshape(rdp.shape), rdp) < TOL


@pytest.mark.parametrize("N,K", [(256, 8), (512, 16), (128, 32), (1024, 8), (700, 5)])
def test_nn_adjoint_scaled_checkpointed(gpu, N, K):
    """Balanced lossy Clements mesh (scaled unitaries): the checkpointed adjoint in one launch
    runs the specialised kernel with checkpoint reloads at stage starts."""
    import meshes_amd._lib as Lb
    X, B = "mzi_generic", 64
    g, p, s, u, rng = _setup(X, "nn", N, N, B, seed=N + K, lossy=True, clem=True)
    mesh = _mesh(g, X, p, s)
    L = mesh.L
    flags = Lb.SCALED_UNITARY
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, X, N, L, B, flags=flags) == "nn"
    ub = torch.as_tensor(u).cuda().contiguous()
    ldp, lds = mesh._ld()
    out = torch.empty_like(ub)
    ck = torch.empty(((L + K - 1) // K, B, N), dtype=ub.dtype, device=ub.device)
    ws, nb = mesh._workspace(Lb.OP_FWDPROP, B)
    Lb.run(Lb.OP_FWDPROP, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, None, ldp, mesh.p_splitter, lds,
           ub, None, out, None, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
    gout = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().contiguous()
    res = {}
    for fl in (flags, flags | Lb.GENERIC_ADJOINT):
        dp = torch.zeros_like(mesh.p_phase)
        uin, gin = torch.empty_like(ub), torch.empty_like(ub)
        _call(mesh, Lb.OP_BACKDIFF, B, out, gout, uin, gin, dp, None, None, fl, ckpt=ck, K=K)
        res[fl] = (dp, uin, gin)
    torch.cuda.synchronize()
    if _same_shape("mzi", N, L, B, None):  # scaled unitaries take the MZI plan
        for a, b in zip(res[flags], res[flags | Lb.GENERIC_ADJOINT]):
            assert torch.equal(a, b)
    dp, uin, gin = res[flags]
    rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128),
                              _np(gout).astype(np.complex128), u_in=u.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL
    assert O.rel_l2(_np(uin), u) < TOL


def test_nn_adjoint_not_used_elsewhere(gpu):
    """FFT geometries, real fields and lossy crossings keep the general kernel."""
    import meshes_amd._lib as Lb
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, "mzi", 256, 8, 64, fft=True) == "general"
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, "orth", 256, 8, 64) == "general"
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, "mzi_generic", 256, 8, 64) == "general"
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, "mzi", 1024, 8, 64, fft=True) == "general"
    assert Lb.adjoint_kernel(Lb.OP_BACKDIFF, "orth", 1024, 8, 64) == "general"
    assert Lb.adjoint_kernel(Lb.OP_FWDPROP, "mzi", 256, 8, 64) == "general"


# Staged forward (W = 16, at least two waves per SIMD): odd and even depths, one and two
# vectors per lane, padded widths, real fields, lossy crossings with checkpoints.
FWD = [("mzi", 1024, 13, 2048, None), ("mzi", 1024, 8, 4096, (16, 2, 4)), ("mzi", 1000, 7, 2100, (16, 1, 4)),
       ("sym", 600, 9, 2048, (16, 1, 4)), ("orth", 1024, 6, 2048, (16, 1, 4)),
       ("mzi_generic", 800, 10, 2048, (16, 1, 4)),
       ("mzi", 1024, 1, 2048, None), ("mzi", 1024, 2, 2048, None)]


@pytest.mark.parametrize("case", FWD, ids=[_id(c) for c in FWD])
def test_nn_forward(gpu, case):
    """fwdprop of wide NN meshes runs k_fwd_nn (rows staged through LDS): bitwise equal to the
    general forward and within 1e-5 of the oracle; with checkpoints the stored fields are
    bitwise equal too."""
    import meshes_amd._lib as Lb
    X, N, L, B, tile = case
    t4 = None if tile is None else tuple(tile) + (0,)
    assert Lb.forward_kernel(Lb.OP_FWDPROP, X, N, L, B, tile=t4) == "nn"
    assert Lb.forward_kernel(Lb.OP_FWDPROP, X, N, L, B, tile=t4, flags=Lb.GENERIC_KERNELS) == "general"
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=11 * N + L, lossy=X == "mzi_generic")
    mesh = _mesh(g, X, p, s)
    tile = None if tile is None else tuple(tile) + (0,)
    ub = torch.as_tensor(u).cuda().contiguous()
    K = 4
    res = {}
    for flags in (0, Lb.GENERIC_KERNELS):
        out = torch.empty_like(ub)
        ck = torch.empty(((L + K - 1) // K, B, N), dtype=ub.dtype, device=ub.device)
        ldp, lds = mesh._ld()
        g0, g1 = mesh._geom()
        sp = mesh.p_splitter if mesh.p_splitter.numel() else None
        nb = Lb.workspace_size(Lb.OP_FWDPROP, X, N, L, B, False, tile)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=ub.device)
        Lb.run(Lb.OP_FWDPROP, X, N, L, B, False, g0, g1, mesh.p_phase, None, ldp, sp, lds, ub, None, out, None, N,
               ckpt=ck, ckpt_every=K, tile=tile, workspace=ws, ws_bytes=nb, flags=flags)
        res[flags] = (out, ck)
    torch.cuda.synchronize()
    assert torch.equal(res[0][0], res[Lb.GENERIC_KERNELS][0])
    assert torch.equal(res[0][1], res[Lb.GENERIC_KERNELS][1])
    ref = O.mesh_fwd(g, X, p.astype(np.float64), _f64(s), u.astype(np.complex128 if X != "orth" else np.float64))
    assert O.rel_l2(_np(res[0][0]), ref) < TOL


# Staged forward-mode walk (W >= 8, one vector per wave, a batch that gives every CU a
# workgroup): odd and even depths, one and two layers, complex and real fields.
FWDDIFF = [("mzi", 1024, 9, 2100), ("sym", 700, 12, 2048), ("orth", 1024, 7, 2048), ("mzi", 1024, 1, 2048),
           ("mzi", 1000, 2, 2064), ("mzi_generic", 800, 6, 2048), ("mzi", 512, 11, 4100), ("orth", 300, 8, 2048),
           ("sym", 400, 3, 2048)]


@pytest.mark.parametrize("case", FWDDIFF, ids=["-".join(map(str, c)) for c in FWDDIFF])
def test_nn_fwddiff(gpu, case):
    """fwddiff of wide NN meshes runs k_fwddiff_nn (both row tables staged through LDS):
    bitwise equal to the general fwddiff and within 1e-5 of the oracle's JVP."""
    import meshes_amd._lib as Lb
    X, N, L, B = case
    assert Lb.forward_kernel(Lb.OP_FWDDIFF, X, N, L, B) == "nn"
    assert Lb.forward_kernel(Lb.OP_FWDDIFF, X, N, L, B, flags=Lb.GENERIC_KERNELS) == "general"
    assert Lb.forward_kernel(Lb.OP_FWDDIFF, X, N, L, 256) == "general"
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=13 * N + L, lossy=X == "mzi_generic")
    mesh = _mesh(g, X, p, s)
    dp = rng.standard_normal(p.shape).astype(np.float32)
    du = (rng.standard_normal(u.shape) + (0 if O.REAL_FIELD[X] else 1j * rng.standard_normal(u.shape)))
    du = du.astype(u.dtype)
    ub, dub = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(du).cuda().contiguous()
    dpt = torch.as_tensor(dp.reshape(-1)).cuda()
    g0, g1 = mesh._geom()
    ldp, lds = mesh._ld()
    sp = mesh.p_splitter if mesh.p_splitter.numel() else None
    res = {}
    for flags in (0, Lb.GENERIC_KERNELS):
        out, dout = torch.empty_like(ub), torch.empty_like(ub)
        nb = Lb.workspace_size(Lb.OP_FWDDIFF, X, N, L, B, False)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=ub.device)
        Lb.run(Lb.OP_FWDDIFF, X, N, L, B, False, g0, g1, mesh.p_phase, dpt, ldp, sp, lds, ub, dub, out, dout, N,
               workspace=ws, ws_bytes=nb, flags=flags)
        res[flags] = (out, dout)
    torch.cuda.synchronize()
    assert torch.equal(res[0][0], res[Lb.GENERIC_KERNELS][0])
    assert torch.equal(res[0][1], res[Lb.GENERIC_KERNELS][1])
    ref, dref = O.mesh_jvp(g, X, p.astype(np.float64), _f64(s), u.astype(np.complex128), dp.astype(np.float64),
                           du.astype(np.complex128))
    assert O.rel_l2(_np(res[0][0]), ref) < TOL
    assert O.rel_l2(_np(res[0][1]), dref) < TOL
