"""The CPU-mesh bridge and the phase screen on the GPU.

StructuredMeshNetwork-like objects (attributes as mesh.py:240-290 sets them) are turned into
MeshNetworkGPU by meshes_amd.bridge.to_gpu and checked against the golden vectors of the
compiled reference engine c/dot.cpp (a ButterflyNetwork geometry with its permutations and the
output phase screen; a Clements mesh with the input screen) and against the FP64 oracle at larger
sizes: forward, forward-mode and reverse-mode, 1e-5 relative L2.  MeshOptimizer lowers its
objective; grad_phi_target agrees with the oracle (including its conjugated-input path).
"""
import os

import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))


class MZICrossing:  # the bridge reads the crossing's class name and parameter counts
    n_phase, n_splitter = 2, 2


class Net:
    """A StructuredMeshNetwork-like object (mesh.py:240-290 attributes)."""

    def __init__(self, N, lens, shifts, p_crossing, phi_out, perm, phi_pos, is_phase):
        self.N, self.lens, self.shifts = N, list(lens), list(shifts)
        self.L = len(lens)
        self.X = MZICrossing()
        self.p_phase = np.concatenate([np.asarray(p_crossing).reshape(-1)] + ([np.asarray(phi_out)] if is_phase else []))
        self.p_splitter = 0.0
        self.perm = perm if perm is not None else [None] * (self.L + 1)
        self.phi_pos, self.is_phase = phi_pos, is_phase


def _inds(net):
    from meshes_amd.bridge import strides_from_perms
    N = net.N
    if all(q is None for q in net.perm):
        return np.concatenate([np.arange(n * (N // 2) + s // 2, n * (N // 2) + s // 2 + k)
                               for n, (s, k) in enumerate(zip(net.shifts, net.lens))])
    _, slots = strides_from_perms(N, net.lens, net.shifts, net.perm)
    return np.concatenate([n * (N // 2) + np.asarray(sl) for n, sl in enumerate(slots)])


def _c(a):
    return torch.as_tensor(np.ascontiguousarray(a.astype(np.complex64))).cuda()


@pytest.mark.parametrize("name", ["butterfly16_screen_out", "clements12_screen_in"])
def test_bridge_matches_reference_engine(gpu, name):
    from meshes_amd.bridge import to_gpu
    d = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    N = int(d["N"])
    perms = [np.asarray(q) for q in d["perms"]] if d["perms"].size else None
    net = Net(N, d["lens"], d["shifts"], d["p"], d["ph"], perms, ["in", "out"][int(d["pos"])], True)
    mesh = to_gpu(net)
    u = _c(d["u_in"])
    y = mesh.dot(u.T).T
    assert O.rel_l2(y.cpu().numpy(), d["fwd_out"]) < TOL
    inds = _inds(net)
    dp = np.zeros(mesh.n_phase, np.float32)
    dp_cr = dp[:mesh.n_cr * 2].reshape(-1, 2)
    dp_cr[inds] = d["dp"]
    dp[mesh.n_cr * 2:] = d["dph"]
    y1, dy1 = mesh.dot(u.T, dp=dp)
    assert O.rel_l2(y1.T.cpu().numpy(), d["jvp_out"]) < TOL
    assert O.rel_l2(dy1.T.cpu().numpy(), d["jvp_dout"]) < TOL
    gp, gin = mesh.grad_phi(_c(d["fwd_out"]).T, _c(d["g_out"]).T, vpos="out")
    gp = gp.cpu().numpy()
    ref_cr = O.grad_from_dJdT("mzi", d["p"], None, d["vjp_dJdT"].transpose(2, 0, 1))
    assert O.rel_l2(gp[:mesh.n_cr * 2].reshape(-1, 2)[inds], ref_cr) < TOL
    assert O.rel_l2(gp[mesh.n_cr * 2:], d["vjp_dph"]) < TOL
    assert O.rel_l2(gin.T.cpu().numpy(), d["vjp_g_in"]) < TOL


@pytest.mark.parametrize("pos", ["in", "out"])
def test_phase_screen_at_size(gpu, pos):
    """Butterfly N=256 (255 permuted layers) with a phase screen against the FP64 oracle."""
    from meshes_amd.bridge import to_gpu
    rng = np.random.RandomState(7 if pos == "in" else 8)
    N, B = 256, 48
    lens, shifts, perms = O.butterfly_perms(N)
    p = rng.uniform(0, 2 * np.pi, (sum(lens), 2))
    ph = rng.uniform(0, 2 * np.pi, N)
    net = Net(N, lens, shifts, p, ph, perms, pos, True)
    mesh = to_gpu(net)
    ipos = {"in": 0, "out": 1}[pos]
    p32 = p.astype(np.float32).astype(np.float64)
    ph32 = ph.astype(np.float32).astype(np.float64)
    T = O.crossing_T("mzi", p32).transpose(1, 2, 0)
    u = O.unit_cn(rng, B, N).astype(np.complex64)
    ref = O.structured_run(N, lens, shifts, T, u.astype(complex), 0, ph=ph32, perms=perms, pos=ipos)
    out = mesh.dot(_c(u).T).T.cpu().numpy()
    assert O.rel_l2(out, ref) < TOL
    yt = O.unit_cn(rng, B, N)
    g = 2 * (out.astype(complex) - yt)
    pbk = np.array([np.argsort(q) for q in perms])
    _, gin, dJdT, dph = O.structured_run(N, lens, shifts, T, out.astype(complex), 2, dv=g, ph=ph32, perms=pbk,
                                         pos=ipos)
    gp, gg = mesh.grad_phi(_c(out).T, _c(g).T, vpos="out")
    gp = gp.cpu().numpy()
    ref_cr = O.grad_from_dJdT("mzi", p32, None, dJdT.transpose(2, 0, 1))
    assert O.rel_l2(gp[:mesh.n_cr * 2].reshape(-1, 2)[_inds(net)], ref_cr) < TOL
    assert O.rel_l2(gp[mesh.n_cr * 2:], dph) < TOL
    assert O.rel_l2(gg.T.cpu().numpy(), gin) < TOL


def test_grad_phi_target_and_optimizer(gpu):
    """grad_phi_target (conjugated target input) against the oracle, then MeshOptimizer
    (gpu/opt.py) lowers |offdiag(U_mesh U^H)|^2 toward a target built by another mesh."""
    from meshes_amd.bridge import MeshOptimizer, to_gpu
    rng = np.random.RandomState(9)
    N = 8
    g = O.clements(N)
    lens, shifts = list(g.lens), list(g.shifts)
    n_cr = sum(lens)
    p0 = rng.uniform(0, 2 * np.pi, (n_cr, 2))
    pt = rng.uniform(0, 2 * np.pi, (n_cr, 2))
    net = Net(N, lens, shifts, p0, None, None, "out", False)
    Tt = O.crossing_T("mzi", pt).transpose(1, 2, 0)
    U = O.structured_dense(N, lens, shifts, Tt).astype(np.complex64)
    mesh = to_gpu(net)
    J, dp = mesh.grad_phi_target(torch.as_tensor(U).cuda())
    # oracle: M = U_mesh U^H, V = offdiag(M), J = |V|^2, dJ/dM = 2 V
    T0 = O.crossing_T("mzi", p0.astype(np.float32).astype(np.float64)).transpose(1, 2, 0)
    Uh = U.astype(complex).conj().T
    M = Uh.T @ O.structured_dense(N, lens, shifts, T0).T  # rows: (U_mesh U^H) columns as vectors
    V = M - np.diag(np.diag(M))
    assert np.isclose(J, np.sum(np.abs(V) ** 2), rtol=1e-4)
    _, _, dJdT, _ = O.structured_run(N, lens, shifts, T0, M, 2, dv=2 * V)
    ref = O.grad_from_dJdT("mzi", p0.astype(np.float32).astype(np.float64), None, dJdT.transpose(2, 0, 1))
    assert O.rel_l2(dp.cpu().numpy().reshape(-1, 2)[_inds(net)], ref) < TOL
    opt = MeshOptimizer(net, U)
    res = opt.run("L-BFGS-B", maxiter=60)
    assert res.yt[-1] < 0.5 * res.yt[0], (res.yt[0], res.yt[-1])
