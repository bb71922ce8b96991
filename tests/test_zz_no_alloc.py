"""meshprop_run_noalloc (the library's static device scratch in place of a caller workspace)
against meshprop_run, through the C ABI.

Bar: equal results to the workspace entry point on consecutive calls of falling size (bitwise
where the adjoint's dp partials fit the scratch, since the plan and summation order are then the
same; within 1e-5 relative L2 where the scratch forces fewer adjoint workgroups), and the small
case against the FP64 oracle.  The scratch keeps the previous call's contents, so a stale
partial or table would show up as a mismatch on the smaller second call.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


class Case:
    def __init__(self, N, L, B, seed, clements=False):
        from meshes_amd import MeshNetworkGPU
        rng = np.random.RandomState(seed)
        self.g = g = O.clements(N) if clements else O.random_geometry_nn(rng, N, L)
        self.N, self.L, self.B = N, g.L, B
        self.p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, 2)).astype(np.float32)
        self.mesh = MeshNetworkGPU(N, g.L, g.lens, g.shifts, self.p, np.zeros(0), "mzi", is_phase=False)
        dev = self.mesh.device
        self.x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64), device=dev)
        self.t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64), device=dev)
        self.dy = torch.as_tensor((rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N)))
                                  .astype(np.complex64), device=dev)
        self.dp0 = torch.as_tensor(rng.standard_normal(self.mesh.n_phase).astype(np.float32), device=dev)

    def call(self, op, noalloc):
        """One call of op through meshprop_run (noalloc=False) or meshprop_run_noalloc."""
        import meshes_amd._lib as Lb
        m = self.mesh
        args = dict(target=None, flags=0)
        out = torch.empty_like(self.x)
        gin = torch.empty_like(self.x)
        dp = self.dp0.clone()  # accumulated into (gpu/mesh.py:214 semantics)
        if op == Lb.OP_FWDGRAD:
            u_a, du_a, u_b, du_b = self.x, None, out, gin
            args["target"] = self.t
        elif op == Lb.OP_BACKDIFF:
            u_a, du_a, u_b, du_b = self.y, self.dy, out, gin
        else:
            u_a, du_a, u_b, du_b = self.x, None, out, None
            dp = None
        pos = (op, "mzi", m.N, m.L, self.B, False, m.lens, m.shifts, m.p_phase, dp, m.N, None, 0, u_a, du_a, u_b,
               du_b, m.N)
        if noalloc:
            Lb.run_noalloc(*pos, **args)
        else:
            ws, nb = m._workspace(op, self.B)
            Lb.run(*pos, workspace=ws, ws_bytes=nb, **args)
        torch.cuda.synchronize()
        return out, gin, dp

    @property
    def y(self):
        return self.mesh.dot(self.x.T).T.contiguous()


def _rel(a, b):
    a = a.detach().cpu().numpy().astype(np.complex128)
    b = b.detach().cpu().numpy().astype(np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_noalloc_matches_workspace_entry_on_consecutive_calls(gpu):
    import meshes_amd._lib as Lb
    # 1) the fwd+grad step at N=256 L=256 B=4096 (its partials fit: same plan, bitwise equal)
    print("noalloc: fwd+grad N=256 L=256 B=4096", flush=True)
    big = Case(256, 256, 4096, seed=11, clements=True)
    ra, rb = big.call(Lb.OP_FWDGRAD, False), big.call(Lb.OP_FWDGRAD, True)
    for a, b in zip(ra, rb):
        assert torch.equal(a, b)
    # 2) a smaller backdiff right after (dp accumulated into a pre-filled buffer)
    print("noalloc: backdiff N=64 L=20 B=77", flush=True)
    small = Case(64, 20, 77, seed=12)
    sa, sb = small.call(Lb.OP_BACKDIFF, False), small.call(Lb.OP_BACKDIFF, True)
    for a, b in zip(sa, sb):
        assert torch.equal(a, b)
    # the small case against the FP64 oracle
    y = small.y.cpu().numpy().astype(np.complex128)
    rdp, rgin, _ = O.mesh_vjp(small.g, "mzi", small.p.astype(np.float64), None, y,
                              small.dy.cpu().numpy().astype(np.complex128))
    got = (sb[2] - small.dp0).cpu().numpy().reshape(rdp.shape)
    assert O.rel_l2(got, rdp) < TOL
    assert O.rel_l2(sb[1].cpu().numpy(), rgin) < TOL
    # 3) fwdprop, smaller again
    print("noalloc: fwdprop N=32 L=9 B=5", flush=True)
    tiny = Case(32, 9, 5, seed=13)
    fa, fb = tiny.call(Lb.OP_FWDPROP, False), tiny.call(Lb.OP_FWDPROP, True)
    assert torch.equal(fa[0], fb[0])
    ref = O.mesh_fwd(tiny.g, "mzi", tiny.p.astype(np.float64), None, tiny.x.cpu().numpy().astype(np.complex128))
    assert O.rel_l2(fb[0].cpu().numpy(), ref) < TOL


def test_noalloc_fewer_workgroups_when_partials_do_not_fit(gpu):
    """N=512 L=512 B=2048 fwd+grad: 256 MiB of partials as planned for meshprop_run; the
    no-alloc call runs the adjoint on fewer workgroups, within 1e-5 of the workspace entry."""
    import meshes_amd._lib as Lb
    print("noalloc: fwd+grad N=512 L=512 B=2048 (grid capped)", flush=True)
    c = Case(512, 512, 2048, seed=14, clements=True)
    assert Lb.workspace_size(Lb.OP_FWDGRAD, "mzi", 512, 512, 2048) > Lb.NOALLOC_SCRATCH_BYTES
    ra, rb = c.call(Lb.OP_FWDGRAD, False), c.call(Lb.OP_FWDGRAD, True)
    assert torch.equal(ra[0], rb[0])  # the forward is the same plan
    assert _rel(rb[1], ra[1]) < TOL
    assert _rel(rb[2] - c.dp0, ra[2] - c.dp0) < TOL
    # and straight after, a small call again
    print("noalloc: fwd+grad N=128 L=16 B=40", flush=True)
    s = Case(128, 16, 40, seed=15)
    sa, sb = s.call(Lb.OP_FWDGRAD, False), s.call(Lb.OP_FWDGRAD, True)
    for a, b in zip(sa, sb):
        assert torch.equal(a, b)
