"""GPU parity of the scaled-unitary adjoint (MESHPROP_SCALED_UNITARY) for generic MZI crossings
with balanced arm gains (T = c U, the loss model of the N=512 training workload).

The adjoint then runs as for unitary crossings on the inverse rows alone, with
T^H = |det T| T^-1 and the K-form gradient K = T^-1 dT/dp.  Bars: dp, dJ/du_in and the
reconstructed input within 1e-5 relative L2 of the FP64 oracle's stored-field adjoint
(crossing.py:444-491 generic model, c/dot.cpp:240-278 adjoint), and equal within the same bar
to the general lossy adjoint (two row tables, dT/dp) that the flag replaces.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O
from tests.test_gpu_parity import _f64, _mesh, _np, _setup

pytestmark = pytest.mark.gpu

TOL = 1e-5
X = "mzi_generic"


def _balanced(N, L, B, seed, clem=True):
    g, p, s, u, rng = _setup(X, "nn", N, L, B, seed=seed, lossy=True, clem=clem)
    assert np.all(s[..., 10] == s[..., 11]) and np.all(s[..., 12] == s[..., 13]) and s[0, 0, 10] < 0
    return g, p, s, u, rng


def _fwd_ckpt(mesh, ub, K):
    import meshes_amd._lib as Lb
    N, L, B = mesh.N, mesh.L, ub.shape[0]
    ldp, lds = mesh._ld()
    out = torch.empty_like(ub)
    ck = torch.empty(((L + K - 1) // K, B, N), dtype=ub.dtype, device=ub.device)
    ws, nb = mesh._workspace(Lb.OP_FWDPROP, B)
    Lb.run(Lb.OP_FWDPROP, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, None, ldp, mesh.p_splitter, lds,
           ub, None, out, None, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb)
    return out, ck


def _bwd(mesh, out, gb, ck, K, flags):
    import meshes_amd._lib as Lb
    N, L, B = mesh.N, mesh.L, out.shape[0]
    ldp, lds = mesh._ld()
    dp, uin, gin = torch.zeros_like(mesh.p_phase), torch.empty_like(out), torch.empty_like(out)
    ws, nb = mesh._workspace(Lb.OP_BACKDIFF, B)
    Lb.run(Lb.OP_BACKDIFF, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, dp, ldp, mesh.p_splitter, lds,
           out, gb, uin, gin, N, ckpt=ck, ckpt_every=K, workspace=ws, ws_bytes=nb, flags=flags)
    return dp, uin, gin


@pytest.mark.parametrize("N,K", [(64, 8), (256, 8), (256, 5), (512, 8)])
def test_scaled_checkpointed_adjoint(gpu, N, K):
    """Balanced lossy Clements mesh: checkpointed backdiff through the scaled-unitary adjoint
    (K=8: one launch; K=5: one launch per segment) against the oracle and the general form."""
    import meshes_amd._lib as Lb
    B = 50
    g, p, s, u, rng = _balanced(N, N, B, seed=40 + N)
    mesh = _mesh(g, X, p, s)
    assert mesh._adj_flags(mesh.p_splitter) == Lb.SCALED_UNITARY
    ub = torch.as_tensor(u).cuda().contiguous()
    out, ck = _fwd_ckpt(mesh, ub, K)
    y = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    gb = (2 * (out - y)).contiguous()
    dp, uin, gin = _bwd(mesh, out, gb, ck, K, Lb.SCALED_UNITARY)
    dpg, uing, ging = _bwd(mesh, out, gb, ck, K, 0)
    torch.cuda.synchronize()
    rdp, rgin, _ = O.mesh_vjp(g, X, p.astype(np.float64), _f64(s), _np(out).astype(np.complex128),
                              _np(gb).astype(np.complex128), u_in=u.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL
    assert O.rel_l2(_np(uin), u) < TOL
    assert O.rel_l2(_np(dp), _np(dpg)) < TOL and O.rel_l2(_np(gin), _np(ging)) < TOL


@pytest.mark.parametrize("N,L,B", [(64, 30, 40), (256, 40, 300), (512, 24, 2600), (700, 8, 30)])
def test_scaled_fwdgrad_one_call(gpu, N, L, B):
    """fwd+grad in one call for balanced generic crossings (reconstruction through T^-1): u_out,
    dp and dJ/du_in against the oracle; N=512 B=2600 runs eight-wave W=8 workgroups, N=700 the
    W=16 tile."""
    import meshes_amd._lib as Lb
    g, p, s, u, rng = _balanced(N, L, B, seed=N + L, clem=False)
    mesh = _mesh(g, X, p, s)
    ldp, lds = mesh._ld()
    y = O.unit_cn(rng, B, N).astype(np.complex64)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    out, gin, dp = torch.empty_like(ub), torch.empty_like(ub), torch.zeros_like(mesh.p_phase)
    ws, nb = mesh._workspace(Lb.OP_FWDGRAD, B)
    Lb.run(Lb.OP_FWDGRAD, X, N, L, B, False, mesh.lens, mesh.shifts, mesh.p_phase, dp, ldp, mesh.p_splitter, lds, ub,
           None, out, gin, N, target=yb, workspace=ws, ws_bytes=nb, flags=Lb.SCALED_UNITARY)
    torch.cuda.synchronize()
    p64, s64 = p.astype(np.float64), _f64(s)
    assert O.rel_l2(_np(out), O.mesh_fwd(g, X, p64, s64, u.astype(np.complex128))) < TOL
    gout = 2 * (_np(out).astype(np.complex128) - y)
    rdp, rgin, _ = O.mesh_vjp(g, X, p64, s64, _np(out).astype(np.complex128), gout, u_in=u.astype(np.complex128))
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(gin), rgin) < TOL


def test_scaled_assertion_checked(gpu):
    """An unbalanced crossing under MESHPROP_SCALED_UNITARY makes dp NaN (the prologue's FP64
    check); the mesh object does not assert it for such a mesh and keeps the general adjoint."""
    import meshes_amd._lib as Lb
    N, L, B = 64, 12, 20
    g, p, s, u, rng = _balanced(N, L, B, seed=9, clem=False)
    s[3, :, 11] += 0.1  # unequal arm gains in every crossing of layer 3 (active ones among them)
    assert g.lens[3] > 0
    mesh = _mesh(g, X, p, s)
    assert mesh._adj_flags(mesh.p_splitter) == 0
    ub = torch.as_tensor(u).cuda().contiguous()
    out, ck = _fwd_ckpt(mesh, ub, 4)
    gb = (2 * out).contiguous()
    dp, _, _ = _bwd(mesh, out, gb, ck, 4, Lb.SCALED_UNITARY)
    dpg, _, ging = _bwd(mesh, out, gb, ck, 4, 0)
    torch.cuda.synchronize()
    assert torch.isnan(dp).any()
    assert torch.isfinite(dpg).all() and torch.isfinite(ging).all()


def test_scaled_through_mesh_api(gpu):
    """MeshNetworkGPU.grad_target_l2 and the autograd layer pick the scaled-unitary adjoint for a
    balanced generic mesh; both match the oracle."""
    from meshes_amd.autograd import MeshLayer
    N, L, B = 128, 40, 64
    g, p, s, u, rng = _balanced(N, L, B, seed=5, clem=False)
    mesh = _mesh(g, X, p, s)
    y = O.unit_cn(rng, B, N).astype(np.complex64)
    ub, yb = torch.as_tensor(u).cuda().contiguous(), torch.as_tensor(y).cuda().contiguous()
    J, dp, uo = mesh.grad_target_l2(ub.T, yb.T)
    layer = MeshLayer(mesh, ckpt_every=8)
    out = layer(ub)
    torch.sum(torch.abs(out - yb) ** 2).backward()
    torch.cuda.synchronize()
    p64, s64 = p.astype(np.float64), _f64(s)
    out64 = O.mesh_fwd(g, X, p64, s64, u.astype(np.complex128))
    rdp, _, _ = O.mesh_vjp(g, X, p64, s64, out64, 2 * (out64 - y), u_in=u.astype(np.complex128))
    assert O.rel_l2(_np(uo.T), out64) < TOL
    assert O.rel_l2(_np(dp).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(_np(layer.p.grad).reshape(rdp.shape), rdp) < TOL
