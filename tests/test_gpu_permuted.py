"""Meshes with arbitrary permutations between layers (PermutedMeshGPU, the meshprop_permute
gather between segments of NN layers) against the FP64 oracle's packed-form engine
(structured_run, pinned by golden vectors of the compiled c/dot.cpp with permutations):
forward, forward mode along a random phase direction, and reverse mode (dp and dJ/dv_in,
output and input form), with the phase screen at the input or output, splitter errors, real
fields; 1e-5 relative L2.  The reference GPU path rejects permuted meshes (mesh.py:573); the
CPU engine applies the gathers between layers (c/dot.cpp:150-152, 289-291).
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


class MZICrossing:
    n_phase, n_splitter = 2, 2


class Net:
    """A StructuredMeshNetwork-like object (mesh.py:240-290 attributes) with splitter errors."""

    def __init__(self, N, lens, shifts, p_crossing, s_crossing, phi_out, perm, phi_pos, is_phase):
        self.N, self.lens, self.shifts = N, list(lens), list(shifts)
        self.L = len(lens)
        self.X = MZICrossing()
        self.p_phase = np.concatenate([np.asarray(p_crossing).reshape(-1)] + ([np.asarray(phi_out)] if is_phase else []))
        self.p_splitter = np.asarray(s_crossing)  # (n_cr, n_splitter), as mesh.py:335-341 reads it
        self.perm = perm
        self.phi_pos, self.is_phase = phi_pos, is_phase


def _geometry(rng, N, L, every):
    g = O.random_geometry_nn(rng, N, L)
    lens, shifts = [int(x) for x in g.lens], [int(x) for x in g.shifts]
    perms = [rng.permutation(N) if (n % every == 0 or n == L) else None for n in range(L + 1)]
    perms[1] = rng.permutation(N)  # two gathers in a row around layer 0
    return lens, shifts, perms


def _slots(N, lens, shifts):
    """Slot index (layer * N/2 + m) of every packed crossing."""
    return np.concatenate([n * (N // 2) + s // 2 + np.arange(k) for n, (s, k) in enumerate(zip(shifts, lens))]
                          + [np.zeros(0, np.int64)]).astype(np.int64)


CASES = [("mzi", 64, 20, 48, "out", 3), ("mzi", 64, 21, 33, "in", 4), ("mzi", 128, 9, 40, None, 1),
         ("sym", 128, 33, 24, None, 5), ("orth", 96, 17, 30, None, 3), ("mzi", 512, 64, 256, "out", 8)]


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c)) for c in CASES])
def test_permuted_mesh_against_oracle(gpu, case):
    from meshes_amd.bridge import _grid, to_gpu
    from meshes_amd.permuted import PermutedMeshGPU
    X, N, L, B, pos, every = case
    rng = np.random.RandomState(N + L + every)
    lens, shifts, perms = _geometry(rng, N, L, every)
    nT = sum(lens)
    n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
    p = rng.uniform(0, 2 * np.pi, (nT, n_p)).astype(np.float32)
    s = (0.03 * rng.standard_normal((nT, n_s))).astype(np.float32) if n_s else None
    is_phase = pos is not None
    ph = rng.uniform(0, 2 * np.pi, N).astype(np.float32) if is_phase else None
    if X == "mzi":
        mesh = to_gpu(Net(N, lens, shifts, p, s, ph, perms, pos or "out", is_phase))
        assert isinstance(mesh, PermutedMeshGPU)
    else:
        p_slot = np.nan_to_num(_grid(p, N, lens, shifts, n_p)).astype(np.float32).reshape(-1)
        s_slot = (np.nan_to_num(_grid(s, N, lens, shifts, n_s)).astype(np.float32).reshape(-1) if n_s
                  else np.zeros(0, np.float32))
        mesh = PermutedMeshGPU(N, L, lens, shifts, perms, p_slot, s_slot, X, is_phase=False)
    assert len(mesh.segments) >= 2
    real = O.REAL_FIELD[X]
    ipos = {"in": 0, "out": 1, None: -1}[pos]
    p64, s64 = p.astype(np.float64), None if s is None else s.astype(np.float64)
    ph64 = None if ph is None else ph.astype(np.float64)
    T = O.crossing_T(X, p64, s64).transpose(1, 2, 0)
    dt = torch.float32 if real else torch.complex64
    u = O.unit_cn(rng, B, N, real=real)
    ut = torch.as_tensor(u.astype(np.float32 if real else np.complex64)).cuda()

    full = [np.arange(N) if x is None else np.asarray(x) for x in perms]  # as mesh.py:347 packs them
    # forward
    ref = O.structured_run(N, lens, shifts, T, u, 0, ph=ph64, perms=full, pos=ipos)
    out = mesh.dot(ut.T).T
    assert O.rel_l2(out.cpu().numpy(), ref) < TOL

    # forward mode along a random direction (inactive slots carry none)
    slots = _slots(N, lens, shifts)
    q = rng.standard_normal((nT, n_p))
    q /= np.linalg.norm(q)
    dp = np.zeros(mesh.n_phase, np.float32)
    dp[:mesh.n_cr * n_p].reshape(-1, n_p)[slots] = q
    dph = None
    if is_phase:
        dph = rng.standard_normal(N) / np.sqrt(N)
        dp[mesh.n_cr * n_p:] = dph
    dT = np.einsum("tk,tkij->ijt", q, O.crossing_dT(X, p64, s64))
    rv, rdv = O.structured_run(N, lens, shifts, T, u, 1, dT=dT, ph=ph64, dph=dph, perms=full, pos=ipos)
    y, dy = mesh.dot(ut.T, dp=torch.as_tensor(dp).cuda())
    assert O.rel_l2(y.T.cpu().numpy(), rv) < TOL
    assert O.rel_l2(dy.T.cpu().numpy(), rdv) < TOL

    # reverse mode from the output (inverse permutations, as mesh.py:381 passes them)
    g = O.unit_cn(rng, B, N, real=real)
    pbk = [np.argsort(x) for x in full]
    _, rgin, dJdT, rdph = O.structured_run(N, lens, shifts, T, ref, 2, dv=g, ph=ph64, perms=pbk, pos=ipos)
    rdp = O.grad_from_dJdT(X, p64, s64, dJdT.transpose(2, 0, 1))
    gt = torch.as_tensor(g.astype(np.float32 if real else np.complex64)).cuda()
    dpo, gin = mesh.grad_phi(out.T, gt.T, vpos="out")
    dpo = dpo.cpu().numpy()
    assert O.rel_l2(dpo[:mesh.n_cr * n_p].reshape(-1, n_p)[slots], rdp) < TOL
    if is_phase:
        assert O.rel_l2(dpo[mesh.n_cr * n_p:], rdph) < TOL
    assert O.rel_l2(gin.T.cpu().numpy(), rgin) < TOL
    # input form (the field propagated forward first) and accumulation into a caller's dp
    acc = torch.full((mesh.n_phase,), 0.25, dtype=torch.float32, device="cuda")
    dpi, gin2 = mesh.grad_phi(ut.T, gt.T, out_dp=acc, vpos="in")
    assert dpi is acc
    assert O.rel_l2(dpi.cpu().numpy() - 0.25, dpo) < TOL
    assert O.rel_l2(gin2.T.cpu().numpy(), rgin) < TOL
    assert out.dtype == dt


def test_permute_kernel_gathers(gpu):
    """meshprop_permute: out[b, w] = in[b, perm[w]] for complex64 and float32 fields, ragged
    batches and N beyond one block."""
    import meshes_amd._lib as Lb
    rng = np.random.RandomState(3)
    for N, B, dt in [(6, 1, torch.complex64), (300, 130, torch.complex64), (1024, 65, torch.float32)]:
        perm = rng.permutation(N)
        x = torch.randn(B, N, dtype=dt, device="cuda")
        y = torch.empty_like(x)
        Lb.permute(N, B, torch.as_tensor(perm.astype(np.int32), device="cuda"), x, y, N)
        torch.cuda.synchronize()
        assert torch.equal(y, x[:, torch.as_tensor(perm, device="cuda")])


def test_optimizer_on_permuted_mesh(gpu):
    """MeshOptimizer (gpu/opt.py:16-69) on a mesh with arbitrary permutations: the packed phases
    map onto the NN slots of the segments, and a few L-BFGS steps lower the objective."""
    from meshes_amd.bridge import MeshOptimizer
    rng = np.random.RandomState(11)
    N, L = 16, 16
    lens, shifts = [N // 2 - (n % 2) for n in range(L)], [n % 2 for n in range(L)]
    perms = [None] * (L + 1)
    perms[5], perms[11] = rng.permutation(N), rng.permutation(N)
    p = rng.uniform(0, 2 * np.pi, (sum(lens), 2))
    net = Net(N, lens, shifts, p, np.zeros((sum(lens), 2)), None, perms, "out", False)
    U = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))[0]
    opt = MeshOptimizer(net, U)
    f0 = opt.f(opt.p)
    res = opt.run("L-BFGS-B", maxiter=20)
    assert res.fun < f0
