"""torch.autograd form: gradients of mesh_apply through the hand-written adjoint.

Checked against the FP64 oracle (stored-field adjoint, exact for lossy meshes), against central
finite differences with h = 1e-3 of the FP64 oracle loss on random phases, and across
checkpoint intervals K in {8, 16, 32}; plus a short Adam run toward a Haar-random target with a
lossy, imperfect mesh (the training form of the lossy-mesh configuration, scaled down).
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _lossy(rng, N, loss_db=0.1, sigma=0.02):
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    s = np.zeros((N, N // 2, 14), np.float32)
    s[..., :2] = sigma * rng.standard_normal((N, N // 2, 2))
    # per-arm gain g_ij: two stages per arm, amplitude e^{g/2} each -> loss_db per crossing
    s[..., 10:] = np.float32(np.log(10 ** (-loss_db / 20)))
    return g, p, s


def _mesh(g, X, p, s):
    from meshes_amd import MeshNetworkGPU
    return MeshNetworkGPU(g.N, g.L, g.lens, g.shifts, p, s, X, is_phase=False)


def test_autograd_unitary_matches_oracle(gpu):
    from meshes_amd.autograd import mesh_apply
    rng = np.random.RandomState(41)
    N, B = 64, 96
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    mesh = _mesh(g, "mzi", p, np.zeros(0))
    x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().requires_grad_(True)
    t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    pt = mesh.p_phase.clone().requires_grad_(True)
    y = mesh_apply(mesh, pt, x)
    loss = torch.sum(torch.abs(y - t) ** 2)
    loss.backward()
    yo = O.mesh_fwd(g, "mzi", p.astype(np.float64), None, x.detach().cpu().numpy().astype(np.complex128))
    rdp, rgx, _ = O.mesh_vjp(g, "mzi", p.astype(np.float64), None, yo, 2 * (yo - t.cpu().numpy()))
    assert O.rel_l2(y.detach().cpu().numpy(), yo) < TOL
    assert O.rel_l2(pt.grad.cpu().numpy().reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(x.grad.cpu().numpy(), rgx) < TOL


@pytest.mark.parametrize("K", [8, 16, 32])
def test_autograd_lossy_checkpointed(gpu, K):
    from meshes_amd.autograd import mesh_apply
    rng = np.random.RandomState(43)
    N, B = 64, 128
    g, p, s = _lossy(rng, N)
    mesh = _mesh(g, "mzi_generic", p, s)
    x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    pt = mesh.p_phase.clone().requires_grad_(True)
    y = mesh_apply(mesh, pt, x, ckpt_every=K)
    loss = torch.sum(torch.abs(y - t) ** 2)
    loss.backward()
    xn = x.cpu().numpy().astype(np.complex128)
    s64 = s.astype(np.float64)
    yo = O.mesh_fwd(g, "mzi_generic", p.astype(np.float64), s64, xn)
    rdp, _, _ = O.mesh_vjp(g, "mzi_generic", p.astype(np.float64), s64, yo, 2 * (yo - t.cpu().numpy()), u_in=xn)
    gp = pt.grad.cpu().numpy().reshape(rdp.shape)
    assert O.rel_l2(y.detach().cpu().numpy(), yo) < TOL
    assert O.rel_l2(gp, rdp) < TOL
    # central differences (h = 1e-3) of the FP64 loss on random active phases
    act = np.argwhere(np.repeat(O.active_mask(g)[..., None], 2, axis=2))
    pick = act[rng.choice(len(act), 16, replace=False)]
    tn = t.cpu().numpy()
    h = 1e-3
    for (l, m, k) in pick:
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[l, m, k] += h
        pm[l, m, k] -= h
        Jp = np.sum(np.abs(O.mesh_fwd(g, "mzi_generic", pp, s64, xn) - tn) ** 2)
        Jm = np.sum(np.abs(O.mesh_fwd(g, "mzi_generic", pm, s64, xn) - tn) ** 2)
        fd = (Jp - Jm) / (2 * h)
        assert abs(gp[l, m, k] - fd) <= 1e-4 * max(1.0, abs(fd)), (l, m, k, gp[l, m, k], fd)


def test_adam_training_toward_haar_target(gpu):
    """Adam on the phases of a lossy, imperfect Clements mesh toward a Haar-random unitary."""
    from scipy.stats import unitary_group
    from meshes_amd.autograd import MeshLayer
    rng = np.random.RandomState(6)
    N, B, steps = 32, 256, 60
    g, p, s = _lossy(rng, N)
    layer = MeshLayer(_mesh(g, "mzi_generic", p, s), ckpt_every=8)
    U = unitary_group.rvs(N, random_state=6)
    x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    t = x @ torch.as_tensor(U.T.astype(np.complex64)).cuda()
    opt = torch.optim.Adam(layer.parameters(), lr=0.02)
    losses = []
    for _ in range(steps):
        opt.zero_grad()
        loss = torch.sum(torch.abs(layer(x) - t) ** 2) / B
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], losses[::10]


@pytest.mark.parametrize("K", [8, 16, 32])
def test_c3_full_size_gradient(gpu, K):
    """Lossy, imperfect Clements mesh at full size (N = L = 512; alpha, beta ~ N(0, 0.02),
    0.1 dB per crossing) on a 32-vector slice: the checkpointed adjoint (checkpoints every K
    layers, one launch each way) against the FP64 stored-field adjoint, and, for K = 16, central
    finite differences (h = 1e-3) of the FP64 loss on 64 random active phases."""
    from meshes_amd.autograd import mesh_apply
    rng = np.random.RandomState(6)
    N, B = 512, 32
    g, p, s = _lossy(rng, N)
    mesh = _mesh(g, "mzi_generic", p, s)
    x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    pt = mesh.p_phase.clone().requires_grad_(True)
    y = mesh_apply(mesh, pt, x, ckpt_every=K)
    torch.sum(torch.abs(y - t) ** 2).backward()
    xn, tn = x.cpu().numpy().astype(np.complex128), t.cpu().numpy().astype(np.complex128)
    p64, s64 = p.astype(np.float64), s.astype(np.float64)
    T = O.slot_T(g, "mzi_generic", p64, s64)
    yo = O.fwd(g, T, xn)
    assert O.rel_l2(y.detach().cpu().numpy(), yo) < TOL
    rdp, _, _ = O.mesh_vjp(g, "mzi_generic", p64, s64, yo, 2 * (yo - tn), u_in=xn)
    gp = pt.grad.cpu().numpy().reshape(rdp.shape)
    assert O.rel_l2(gp, rdp) < TOL
    if K != 16:
        return
    act = np.argwhere(np.repeat(O.active_mask(g)[..., None], 2, axis=2))
    pick = act[rng.choice(len(act), 64, replace=False)]
    h = 1e-3
    for (l, m, k) in pick:
        vals = []
        for sgn in (1, -1):
            pp = p64[l, m].copy()
            pp[k] += sgn * h
            Tp = T.copy()
            Tp[l, m] = O.crossing_T("mzi_generic", pp[None], s64[l, m][None])[0]
            vals.append(np.sum(np.abs(O.fwd(g, Tp, xn) - tn) ** 2))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(gp[l, m, k] - fd) <= 1e-4 * max(1.0, abs(fd)), (l, m, k, gp[l, m, k], fd)


def test_autograd_loss_through_conjugate_view(gpu):
    """A fidelity loss built from y.conj(), J = sum_b |<y_b, t_b>|^2, hands the backward a lazy
    conjugate view; its gradients match the FP64 oracle (meshes convention g = 2 dJ/dy* =
    2 conj(c_b) t_b with c_b = sum_n conj(y_bn) t_bn)."""
    from meshes_amd.autograd import mesh_apply
    rng = np.random.RandomState(47)
    N, B = 32, 40
    g = O.clements(N)
    p = rng.uniform(0, 2 * np.pi, (N, N // 2, 2)).astype(np.float32)
    mesh = _mesh(g, "mzi", p, np.zeros(0))
    x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda().requires_grad_(True)
    t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
    pt = mesh.p_phase.clone().requires_grad_(True)
    y = mesh_apply(mesh, pt, x)
    loss = torch.sum(torch.abs(torch.sum(y.conj() * t, dim=1)) ** 2)
    loss.backward()
    yo = O.mesh_fwd(g, "mzi", p.astype(np.float64), None, x.detach().cpu().numpy().astype(np.complex128))
    tn = t.cpu().numpy().astype(np.complex128)
    c = np.sum(np.conj(yo) * tn, axis=1)
    rdp, rgx, _ = O.mesh_vjp(g, "mzi", p.astype(np.float64), None, yo, 2 * np.conj(c)[:, None] * tn)
    assert O.rel_l2(pt.grad.cpu().numpy().reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(x.grad.cpu().numpy(), rgx) < TOL


def test_autograd_rejects_bad_inputs(gpu):
    """Shapes, sizes and devices are checked before any pointer reaches a kernel."""
    from meshes_amd.autograd import mesh_apply
    rng = np.random.RandomState(48)
    g = O.clements(16)
    mesh = _mesh(g, "mzi", rng.uniform(0, 6, (16, 8, 2)).astype(np.float32), np.zeros(0))
    x = torch.as_tensor(O.unit_cn(rng, 5, 16).astype(np.complex64))
    p = mesh.p_phase.clone()
    with pytest.raises(ValueError):
        mesh_apply(mesh, p, x)  # host tensor
    with pytest.raises(ValueError):
        mesh_apply(mesh, p, x.cuda().T.contiguous())  # (N, B) instead of (B, N)
    with pytest.raises(ValueError):
        mesh_apply(mesh, p[:-1], x.cuda())  # wrong number of phases
    y = mesh_apply(mesh, p.reshape(16, 8, 2).requires_grad_(True), x.cuda())  # any shape of n_phase values
    assert y.shape == (5, 16)
