"""GPU parity of the one-call fwd+grad step (MESHPROP_OP_FWDGRAD) through the C ABI.

One call builds every table in one prologue, propagates u_in forward through every layer and
back through the adjoint of J = sum |u_out - target|^2 (gpu/mesh.py:208-234 with the loss of
gpu/test.py:131).  Bars: u_out, dp and dJ/du_in within 1e-5 relative L2 of the FP64 oracle,
and bitwise equal to separate fwdprop and backdiff calls with the same target at the same
adjoint tile, since both run the same arithmetic and the same fixed-order dp reduction.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O
from tests.test_gpu_parity import _f64, _mesh, _np, _setup

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [("mzi", "nn", 64, 20, 70), ("mzi", "nn", 250, 9, 5), ("mzi", "nn", 6, 5, 1), ("mzi", "fft", 256, 10, 21),
         ("mzi", "nn", 512, 6, 20), ("mzi", "nn", 1000, 4, 9), ("sym", "nn", 200, 8, 17), ("sym", "fft", 128, 7, 19),
         ("orth", "nn", 38, 9, 13), ("orth", "fft", 512, 9, 8)]


def _id(c):
    return "-".join(map(str, c))


def _run(mesh, op, B, u_a, u_b, du_b, dp, target, flags=0, tile=None):
    import meshes_amd._lib as Lb
    ldp, lds = mesh._ld()
    g0, g1 = mesh._geom()
    s = mesh.p_splitter if mesh.p_splitter.numel() else None
    nb = Lb.workspace_size(op, mesh.X, mesh.N, mesh.L, B, mesh.fft, tile)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=u_a.device)
    Lb.run(op, mesh.X, mesh.N, mesh.L, B, mesh.fft, g0, g1, mesh.p_phase, dp, ldp, s, lds, u_a, None, u_b, du_b,
           mesh.N, target=target, tile=tile, workspace=ws, ws_bytes=nb, flags=flags)


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_fwdgrad_oracle_and_two_launch_path(gpu, case):
    import meshes_amd._lib as Lb
    X, kind, N, L, B = case
    g, p, s, u, rng = _setup(X, kind, N, L, B, seed=11 + N + L)
    mesh = _mesh(g, X, p, s)
    y = O.unit_cn(rng, B, N, real=O.REAL_FIELD[X]).astype(u.dtype)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    # one call
    out_f, gin_f = torch.empty_like(ub), torch.empty_like(ub)
    dp_f = torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, B, ub, out_f, gin_f, dp_f, yb)
    # separate fwdprop and backdiff calls at the one call's adjoint tile
    tile = Lb.resolve_tile(Lb.OP_FWDGRAD, X, N, B)
    out_s, rec, gin_s = torch.empty_like(ub), torch.empty_like(ub), torch.empty_like(ub)
    dp_s = torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDPROP, B, ub, out_s, None, None, None)
    _run(mesh, Lb.OP_BACKDIFF, B, out_s, rec, gin_s, dp_s, yb, tile=tile)
    torch.cuda.synchronize()
    assert torch.equal(out_f, out_s)
    assert torch.equal(gin_f, gin_s)
    assert torch.equal(dp_f, dp_s)
    # FP64 oracle
    p64, s64 = p.astype(np.float64), _f64(s)
    ro = O.mesh_fwd(g, X, p64, s64, u.astype(np.complex128))
    assert O.rel_l2(_np(out_f), ro) < TOL
    gout = 2 * (_np(out_f).astype(np.complex128) - y)
    rdp, rgin, _ = O.mesh_vjp(g, X, p64, s64, _np(out_f).astype(np.complex128), gout)
    assert O.rel_l2(_np(dp_f).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin_f), rgin) < TOL


def test_fwdgrad_assign_accumulate_and_in_place(gpu):
    """dp accumulates by default and is overwritten with MESHPROP_DP_ASSIGN; u_out may alias
    u_in, and dJ/du_in is optional."""
    import meshes_amd._lib as Lb
    g, p, s, u, rng = _setup("mzi", "nn", 128, 30, 45, seed=3)
    mesh = _mesh(g, "mzi", p, s)
    y = torch.as_tensor(O.unit_cn(rng, 45, 128).astype(np.complex64)).cuda()
    ub = torch.as_tensor(u).cuda().contiguous()
    out = torch.empty_like(ub)
    dp0 = torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, 45, ub, out, None, dp0, y)
    ones = torch.ones_like(mesh.p_phase)
    dp1 = ones.clone()
    _run(mesh, Lb.OP_FWDGRAD, 45, ub, out, None, dp1, y)
    dp2 = torch.full_like(mesh.p_phase, 7.0)
    _run(mesh, Lb.OP_FWDGRAD, 45, ub, out, None, dp2, y, flags=Lb.DP_ASSIGN)
    inplace = ub.clone()
    dp3 = torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, 45, inplace, inplace, None, dp3, y)
    torch.cuda.synchronize()
    assert torch.equal(dp1, dp0 + ones)
    assert torch.equal(dp2, dp0)
    assert torch.equal(inplace, out) and torch.equal(dp3, dp0)


def test_fwdgrad_c2_full_size(gpu):
    """Config C2 (Clements MZI N=256 L=256 B=4096): the one-call step equals separate fwdprop
    and backdiff calls bitwise, also when its stages are issued one by one (as the benchmark
    times them), and grad_target_l2 (which runs it) returns J and dp of the same loss."""
    import meshes_amd._lib as Lb
    N = 256
    rng = np.random.RandomState(2)
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    pt = np.random.RandomState(3).uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    u = O.unit_cn(np.random.RandomState(4), 4096, N).astype(np.complex64)
    mesh, target = _mesh(g, "mzi", p, None), _mesh(g, "mzi", pt, None)
    ub = torch.as_tensor(u).cuda().contiguous()
    yb = target.dot(ub.T).T.contiguous()
    out_f, dp_f = torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, 4096, ub, out_f, None, dp_f, yb)
    tile = Lb.resolve_tile(Lb.OP_FWDGRAD, "mzi", N, 4096)
    out_s, rec, dp_s = torch.empty_like(ub), torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDPROP, 4096, ub, out_s, None, None, None)
    _run(mesh, Lb.OP_BACKDIFF, 4096, out_s, rec, None, dp_s, yb, tile=tile)
    J, dp_m, uo = mesh.grad_target_l2(ub.T, yb.T)
    out_k, dp_k = torch.empty_like(ub), torch.full_like(mesh.p_phase, 3.0)
    nb = Lb.workspace_size(Lb.OP_FWDGRAD, "mzi", N, N, 4096)
    ws = torch.empty(nb, dtype=torch.uint8, device=ub.device)
    for fl in (Lb.ONLY_PROLOGUE, Lb.SKIP_PROLOGUE | Lb.ONLY_FORWARD,
               Lb.SKIP_PROLOGUE | Lb.SKIP_FORWARD | Lb.SKIP_REDUCE, Lb.ONLY_REDUCE | Lb.DP_ASSIGN):
        Lb.run(Lb.OP_FWDGRAD, "mzi", N, N, 4096, False, mesh.lens, mesh.shifts, mesh.p_phase, dp_k, N, None, 0, ub,
               None, out_k, None, N, target=yb, workspace=ws, ws_bytes=nb, flags=fl)
    torch.cuda.synchronize()
    assert torch.equal(out_k, out_f) and torch.equal(dp_k, dp_f)
    assert torch.equal(out_f, out_s) and torch.equal(dp_f, dp_s)
    assert torch.equal(dp_m, dp_f) and torch.equal(uo, out_f.T)
    assert abs(J.item() - torch.sum(torch.abs(out_f - yb) ** 2).item()) <= 1e-6 * J.item()


@pytest.mark.parametrize("X,N,L,B", [("mzi", 256, 24, 300), ("mzi", 128, 17, 1000), ("mzi", 200, 9, 77),
                                     ("orth", 256, 12, 300), ("sym", 256, 10, 200)])
def test_fwdgrad_eight_wave_workgroups(gpu, X, N, L, B):
    """Adjoint workgroups of eight waves (pipelined stage, one workgroup per CU): same output,
    dp and dJ/du_in as the four-wave default within the oracle bar (different partial sums);
    complex and real (orth) fields."""
    import meshes_amd._lib as Lb
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=N + L)
    mesh = _mesh(g, X, p, s)
    y = O.unit_cn(rng, B, N, real=O.REAL_FIELD[X]).astype(u.dtype)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    w = Lb.resolve_tile(Lb.OP_FWDGRAD, X, N, B)[0]
    assert Lb.resolve_tile(Lb.OP_FWDGRAD, X, N, B, (w, 0, 8, 0))[2] == 8
    res = []
    for tile in (None, (w, 0, 8, 0)):
        out, gin, dp = torch.empty_like(ub), torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
        _run(mesh, Lb.OP_FWDGRAD, B, ub, out, gin, dp, yb, tile=tile)
        res.append((out, gin, dp))
    torch.cuda.synchronize()
    assert torch.equal(res[0][0], res[1][0])
    p64, s64 = p.astype(np.float64), _f64(s)
    ft = np.float64 if O.REAL_FIELD[X] else np.complex128
    gout = 2 * (_np(res[1][0]).astype(ft) - y)
    rdp, rgin, _ = O.mesh_vjp(g, X, p64, s64, _np(res[1][0]).astype(ft), gout)
    for out, gin, dp in res:
        assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
        assert O.rel_l2(_np(gin), rgin) < TOL
    assert O.rel_l2(_np(res[0][2]), _np(res[1][2])) < 1e-6


def test_fwdgrad_several_tiles_per_workgroup(gpu):
    """More vector tiles than adjoint workgroups (B=9000 at N=256: 1125 tiles over at most 512
    workgroups), so workgroups add later tiles into their dp partials; against the oracle."""
    import meshes_amd._lib as Lb
    N, L, B = 256, 16, 9000
    g, p, s, u, rng = _setup("mzi", "nn", N, L, B, seed=77)
    mesh = _mesh(g, "mzi", p, s)
    y = O.unit_cn(rng, B, N).astype(np.complex64)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    out, gin, dp = torch.empty_like(ub), torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, B, ub, out, gin, dp, yb)
    torch.cuda.synchronize()
    p64, s64 = p.astype(np.float64), _f64(s)
    gout = 2 * (_np(out).astype(np.complex128) - y)
    rdp, rgin, _ = O.mesh_vjp(g, "mzi", p64, s64, _np(out).astype(np.complex128), gout)
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL


@pytest.mark.parametrize("tile", [None, (8, 1, 4, 0), (8, 1, 8, 0)], ids=["auto", "w8-4waves", "w8-8waves"])
def test_fwdgrad_w8_pipelined(gpu, tile):
    """N=512 (W=8) unitary adjoint on the pipelined per-crossing stage, four- and eight-wave
    workgroups (B=2600: the automatic choice is eight waves), against the FP64 oracle."""
    import meshes_amd._lib as Lb
    N, L, B = 512, 20, 2600
    g, p, s, u, rng = _setup("mzi", "nn", N, L, B, seed=512)
    mesh = _mesh(g, "mzi", p, s)
    y = O.unit_cn(rng, B, N).astype(np.complex64)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    t = Lb.resolve_tile(Lb.OP_FWDGRAD, "mzi", N, B, tile)
    assert t[0] == 8 and t[1] == 1 and t[2] == (4 if tile and tile[2] == 4 else 8)
    out, gin, dp = torch.empty_like(ub), torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
    _run(mesh, Lb.OP_FWDGRAD, B, ub, out, gin, dp, yb, tile=tile)
    torch.cuda.synchronize()
    p64, s64 = p.astype(np.float64), _f64(s)
    assert O.rel_l2(_np(out), O.mesh_fwd(g, "mzi", p64, s64, u.astype(np.complex128))) < TOL
    gout = 2 * (_np(out).astype(np.complex128) - y)
    rdp, rgin, _ = O.mesh_vjp(g, "mzi", p64, s64, _np(out).astype(np.complex128), gout)
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL
