"""Grouped launch (many meshes, one persistent kernel) against the FP64 oracle and the plain path.

Bar: every job's outputs and gradients within 1e-5 relative L2 of the FP64 oracle; each job's
results bitwise identical alone and inside a mix; hipGraph replays identical to eager launches
(several replays, with eager work and synchronisation between them); the per-mesh plain
launches (meshprop_run, gpu/mesh.py:158-231 call pattern) agree within 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _geom(rng, topo, N):
    if topo == "clements":
        return O.clements(N)
    if topo == "butterfly":
        return O.Geometry(N, int(np.log2(N)), strides=np.arange(int(np.log2(N))))
    return O.random_geometry_nn(rng, N, int(rng.randint(1, 12)))


def _mesh(g, X, p, s):
    from meshes_amd import MeshNetworkGPU
    kw = dict(strides=g.strides) if g.kind == "fft" else {}
    return MeshNetworkGPU(g.N, g.L, g.lens, g.shifts, p, s if s is not None else np.zeros(0), X,
                          is_phase=False, **kw)


def _c64(a):
    return torch.as_tensor(np.ascontiguousarray(a.astype(np.complex64))).cuda()


def _field(a, real):
    """Device field: complex64, or float32 for real (orth) meshes."""
    return torch.as_tensor(np.ascontiguousarray(a.astype(np.float32 if real else np.complex64))).cuda()


class Spec:
    """One job with its FP64 reference."""

    def __init__(self, seed, X, topo, N, B, op, loss="target", gx=False, errors=False):
        rng = np.random.RandomState(seed)
        self.g = g = _geom(rng, topo, N)
        self.X, self.op, self.loss = X, op, loss
        n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
        self.p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
        self.s = (0.05 * rng.standard_normal((g.L, N // 2, n_s))).astype(np.float32) if (errors and n_s) else None
        self.real = real = O.REAL_FIELD[X]
        fdt = np.float32 if real else np.complex64
        self.x = O.unit_cn(rng, B, N, real=real).astype(fdt)
        self.t = O.unit_cn(rng, B, N, real=real).astype(fdt)
        dy = rng.standard_normal((B, N))
        self.dy = (dy if real else dy + 1j * rng.standard_normal((B, N))).astype(fdt)
        self.want_gx = gx
        self.mesh = _mesh(g, X, self.p, self.s)

    def job(self, **kw):
        from meshes_amd.batched import GroupJob
        x = _field(self.x, self.real)
        if self.op == "fwd":
            return GroupJob(self.mesh, "fwd", x=x, **kw)
        gx = torch.empty_like(x) if self.want_gx else None
        tgt = _field(self.t, self.real) if self.loss == "target" else None
        dy = _field(self.dy, self.real) if self.loss == "dy" else None
        if self.op == "grad":
            return GroupJob(self.mesh, "grad", x=x, target=tgt, dy=dy, gx=gx, **kw)
        # backdiff: y given (the forward output), x receives the reconstruction
        y = _field(self.ref_y(), self.real)
        return GroupJob(self.mesh, "backdiff", x=torch.empty_like(x), y=y, target=tgt, dy=dy, gx=gx, **kw)

    def ref_y(self):
        s = None if self.s is None else self.s.astype(np.float64)
        return O.mesh_fwd(self.g, self.X, self.p.astype(np.float64), s, self.x.astype(np.complex128))

    def check(self, job):
        s = None if self.s is None else self.s.astype(np.float64)
        y64 = self.ref_y()
        if self.op == "backdiff":
            y_used = self.ref_y().astype(np.complex64).astype(np.complex128)
            assert O.rel_l2(job.x.cpu().numpy(), self.x) < TOL  # reconstruction
        else:
            assert O.rel_l2(job.y.cpu().numpy(), y64) < TOL
            y_used = job.y.cpu().numpy().astype(np.complex128)
        if self.op == "fwd":
            return
        gout = 2 * (y_used - self.t.astype(np.complex128)) if self.loss == "target" else self.dy.astype(np.complex128)
        rdp, rgx, _ = O.mesh_vjp(self.g, self.X, self.p.astype(np.float64), s, y_used, gout)
        assert O.rel_l2(job.dp.cpu().numpy().reshape(rdp.shape), rdp) < TOL
        if self.want_gx:
            assert O.rel_l2(job.gx.cpu().numpy(), rgx) < TOL


MIX = [
    ("mzi", "clements", 8, 5, "grad"), ("mzi", "clements", 32, 64, "grad"), ("mzi", "clements", 64, 3, "fwd"),
    ("mzi", "clements", 130, 17, "grad"), ("mzi", "clements", 256, 40, "grad"), ("mzi", "clements", 512, 9, "grad"),
    ("mzi", "butterfly", 16, 100, "grad"), ("mzi", "butterfly", 64, 33, "fwd"), ("mzi", "butterfly", 256, 20, "grad"),
    ("sym", "nn", 48, 12, "grad"), ("mzi", "nn", 200, 7, "backdiff"), ("mzi_generic", "nn", 96, 11, "fwd"),
    ("mzi", "clements", 1024, 2, "grad"), ("mzi", "butterfly", 1024, 6, "grad"),
    ("mzi", "nn", 700, 9, "backdiff"), ("sym", "clements", 1024, 5, "grad"),
    ("orth", "clements", 128, 9, "grad"), ("orth", "nn", 700, 5, "grad"), ("orth", "butterfly", 64, 8, "fwd"),
    ("orth", "butterfly", 1024, 3, "grad"), ("orth", "nn", 300, 6, "backdiff"),
]


def _specs():
    out = []
    for i, (X, topo, N, B, op) in enumerate(MIX):
        loss = "dy" if i % 3 == 1 else "target"
        out.append(Spec(100 + i, X, topo, N, B, op, loss=loss, gx=(i % 2 == 0), errors=(X != "mzi")))
    return out


def test_group_matches_oracle(gpu):
    from meshes_amd.batched import MeshGroup
    specs = _specs()
    jobs = [s.job() for s in specs]
    grp = MeshGroup(jobs)
    grp.run()
    torch.cuda.synchronize()
    for sp, j in zip(specs, jobs):
        sp.check(j)


NN_TILES = [(100, (2, 1)), (100, (2, 2)), (120, (2, 4)), (200, (4, 1)), (250, (4, 2)), (400, (8, 1)),
            (64, (2, 1))]


@pytest.mark.parametrize("N,wv", NN_TILES, ids=[f"N{n}-W{w}V{v}" for n, (w, v) in NN_TILES])
def test_group_nn_adjoint_tiles(gpu, N, wv):
    """Every family-0 tile of the NN-specialised adjoint in a grouped launch (NN meshes with one
    64-lane vector group; N = 64 at W = 2 has LPV = 32 and runs the general adjoint) against the
    FP64 oracle, with splitter errors, both losses and dJ/dx."""
    from meshes_amd.batched import MeshGroup
    w, v = wv
    specs = [Spec(7 * N + v, "mzi", "clements", N, 150, "grad", loss="target", gx=True, errors=True),
             Spec(7 * N + v + 1, "sym", "nn", N, 97, "grad", loss="dy", gx=True, errors=True)]
    jobs = [sp.job(tile=(w, v, 0, 0)) for sp in specs]
    grp = MeshGroup(jobs)
    grp.run()
    torch.cuda.synchronize()
    for i, (sp, j) in enumerate(zip(specs, jobs)):
        tile, _ = grp.info(i)
        assert tuple(tile[:2]) == (w, v), tile
        sp.check(j)


def test_group_alone_equals_mix(gpu):
    from meshes_amd.batched import MeshGroup
    specs = _specs()
    mix = [s.job() for s in specs]
    MeshGroup(mix).run()
    torch.cuda.synchronize()
    for sp, jm in zip(specs, mix):
        ja = sp.job()
        MeshGroup([ja]).run()
        torch.cuda.synchronize()
        for name in ("y", "dp", "gx", "x"):
            a, m = getattr(ja, name), getattr(jm, name)
            if a is not None and m is not None:
                assert torch.equal(a, m), (sp.X, sp.g.N, sp.op, name)


def test_group_graph_replay_equals_eager(gpu):
    from meshes_amd.batched import MeshGroup
    specs = _specs()
    jobs = [s.job() for s in specs]
    grp = MeshGroup(jobs)
    grp.run()
    torch.cuda.synchronize()
    eager = [(j.y.clone(), None if j.dp is None else j.dp.clone()) for j in jobs]
    grp.capture()
    for r in range(3):
        for j in jobs:  # eager work between replays: clobber the outputs
            if j.op != "backdiff":  # (backdiff reads y)
                j.y.fill_(float("nan"))
            if j.dp is not None:
                j.dp.fill_(7.0)
        torch.cuda.synchronize()
        grp.run()
        torch.cuda.synchronize()
        for (y0, dp0), j, sp in zip(eager, jobs, specs):
            if sp.op != "backdiff":
                assert torch.equal(j.y, y0), (r, sp.g.N, sp.op)
            if dp0 is not None:
                assert torch.equal(j.dp, dp0), (r, sp.g.N, sp.op)


def test_group_replan_reuses_graph(gpu):
    """A new plan uploaded into the same workspace replays the captured graph correctly."""
    from meshes_amd.batched import MeshGroup
    a = [Spec(7, "mzi", "clements", 64, 50, "grad"), Spec(8, "mzi", "butterfly", 128, 30, "grad")]
    b = [Spec(9, "mzi", "clements", 32, 20, "grad"), Spec(10, "mzi", "clements", 128, 9, "fwd"),
         Spec(11, "sym", "nn", 40, 70, "grad")]
    ja = [s.job() for s in a]
    grp = MeshGroup(ja, workspace=torch.empty(64 << 20, dtype=torch.uint8, device="cuda"))
    grp.capture()
    grp.run()
    torch.cuda.synchronize()
    for sp, j in zip(a, ja):
        sp.check(j)
    graph = grp.graph
    assert graph is not None
    jb = [s.job() for s in b]
    grp.replan(jb)
    assert grp.graph is graph
    grp.run()
    torch.cuda.synchronize()
    for sp, j in zip(b, jb):
        sp.check(j)


def test_group_matches_plain_launches(gpu):
    from meshes_amd.batched import MeshGroup, run_plain
    specs = [s for s in _specs() if s.op != "backdiff"]
    jg = [s.job() for s in specs]
    jp = [s.job() for s in specs]
    MeshGroup(jg).run()
    run_plain(jp)
    torch.cuda.synchronize()
    for a, b in zip(jg, jp):
        assert O.rel_l2(a.y.cpu().numpy(), b.y.cpu().numpy()) < TOL
        if a.dp is not None:
            assert O.rel_l2(a.dp.cpu().numpy(), b.dp.cpu().numpy()) < TOL


def test_group_edges_and_accumulate(gpu):
    from meshes_amd.batched import GroupJob, MeshGroup
    s1 = Spec(21, "mzi", "clements", 64, 1, "grad", gx=True)
    s2 = Spec(22, "mzi", "nn", 30, 45, "grad")
    j1 = s1.job()
    j2 = s2.job(accumulate=True)
    j2.dp.fill_(0.0)
    # an empty batch and a zero-depth mesh ride along
    from meshes_amd import MeshNetworkGPU
    m0 = MeshNetworkGPU(16, 0, np.zeros(0), np.zeros(0), np.zeros(0, np.float32), np.zeros(0), "mzi",
                        is_phase=False)
    x0 = _c64(O.unit_cn(np.random.RandomState(3), 9, 16))
    j0 = GroupJob(m0, "fwd", x=x0)
    je = GroupJob(s2.mesh, "grad", x=torch.empty((0, 30), dtype=torch.complex64, device="cuda"),
                  target=torch.empty((0, 30), dtype=torch.complex64, device="cuda"))
    grp = MeshGroup([j1, j2, j0, je])
    grp.run()
    torch.cuda.synchronize()
    s1.check(j1)
    s2.check(j2)
    assert torch.equal(j0.y, x0)
    assert torch.count_nonzero(je.dp) == 0
    dp1 = j2.dp.clone()
    grp.run()  # accumulate: dp doubles
    torch.cuda.synchronize()
    assert O.rel_l2(j2.dp.cpu().numpy(), 2 * dp1.cpu().numpy()) < 1e-6


def test_group_resident_tables(gpu):
    from meshes_amd.batched import MeshGroup, tables_for
    specs = [Spec(31, "mzi", "clements", 128, 20, "grad"), Spec(32, "mzi", "butterfly", 512, 6, "fwd"),
             Spec(33, "mzi", "clements", 1024, 4, "grad")]
    tabs = [tables_for(s.mesh) for s in specs]
    jobs = [s.job(tables=t, build_tables=True) for s, t in zip(specs, tabs)]
    MeshGroup(jobs).run()
    torch.cuda.synchronize()
    ref = [(j.y.clone(), None if j.dp is None else j.dp.clone()) for j in jobs]
    jobs2 = [s.job(tables=t, build_tables=False) for s, t in zip(specs, tabs)]
    MeshGroup(jobs2).run()
    torch.cuda.synchronize()
    for (y0, dp0), j, sp in zip(ref, jobs2, specs):
        sp.check(j)
        assert torch.equal(j.y, y0)
        if dp0 is not None:
            assert torch.equal(j.dp, dp0)
