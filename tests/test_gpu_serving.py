"""Serving loop: continuous batching over resident meshes.

Each request's result is the same served in the mix as computed alone (bitwise: a request's
arithmetic depends on its own mesh and batch only), agrees with the sequential plain launches
and with the FP64 oracle within 1e-5 relative L2; admission never exceeds the HBM budget.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
SIZES = [16, 24, 40, 64, 100, 130, 256, 300, 512, 700, 1024]


def _meshes(seed=3):
    from meshes_amd import MeshNetworkGPU, geometry
    rng = np.random.RandomState(seed)
    out = []
    for N in SIZES:
        lens, shifts = geometry.clements(N)
        out.append(MeshNetworkGPU(N, N, lens, shifts, geometry.uniform_phases(rng, N, N), np.zeros(0), "mzi",
                                  is_phase=False))
    return out


def _clone(reqs):
    from meshes_amd.serving import Request
    return [Request(r.rid, r.mesh_id, r.op, r.x, r.target, r.t_arrive) for r in reqs]


def test_served_equals_alone_plain_and_oracle(gpu):
    from meshes_amd.batched import GroupJob, MeshGroup
    from meshes_amd.serving import MeshServer, poisson_requests
    meshes = _meshes()
    server = MeshServer(meshes, hbm_budget=3 << 30, slots=3, direct=False)  # every batch grouped: bitwise claim
    reqs = poisson_requests(meshes, rate=50000.0, n=80, seed=5, b_max=300)
    done, st = server.serve(reqs)
    assert st["requests"] == 80 and st["mean_batch"] > 1.0
    seq, _ = server.serve_sequential(_clone(reqs))
    for r, q in zip(done, seq):
        m = meshes[r.mesh_id]
        if r.op == "grad":
            job = GroupJob(m, "grad", x=r.x, target=r.target, tables=server.tables[r.mesh_id], build_tables=False)
        else:
            job = GroupJob(m, "fwd", x=r.x, tables=server.tables[r.mesh_id], build_tables=False)
        MeshGroup([job]).run()
        torch.cuda.synchronize()
        assert torch.equal(job.y, r.y), (r.rid, m.N, r.op)
        assert O.rel_l2(q.y.cpu().numpy(), r.y.cpu().numpy()) < TOL
        if r.op == "grad":
            assert torch.equal(job.dp, r.dp), (r.rid, m.N)
            assert O.rel_l2(q.dp.cpu().numpy(), r.dp.cpu().numpy()) < TOL
    # FP64 oracle on the smaller meshes
    checked = 0
    for r in done:
        m = meshes[r.mesh_id]
        if m.N > 130:
            continue
        g = O.clements(m.N)
        p = m.p_phase.cpu().numpy().astype(np.float64).reshape(m.N, m.N // 2, 2)
        x = r.x.cpu().numpy().astype(np.complex128)
        y = O.mesh_fwd(g, "mzi", p, None, x)
        assert O.rel_l2(r.y.cpu().numpy(), y) < TOL
        if r.op == "grad":
            yg = r.y.cpu().numpy().astype(np.complex128)
            rdp, _, _ = O.mesh_vjp(g, "mzi", p, None, yg, 2 * (yg - r.target.cpu().numpy().astype(np.complex128)))
            assert O.rel_l2(r.dp.cpu().numpy().reshape(rdp.shape), rdp) < TOL
        checked += 1
    assert checked >= 10


def test_served_orth_and_fft_meshes(gpu):
    """Real-field (orth) meshes and FFT meshes beside NN ones: every kernel family of the grouped
    launch (N <= 512; wider NN meshes of MZI crossings; the other wide meshes) in the captured
    graphs of the slots.  Served results equal the same request alone (bitwise), the sequential
    plain launches and, on the small orth meshes, the FP64 oracle (1e-5)."""
    from meshes_amd import MeshNetworkGPU, geometry
    from meshes_amd.batched import GroupJob, MeshGroup
    from meshes_amd.serving import MeshServer, poisson_requests
    rng = np.random.RandomState(12)
    meshes, geoms = [], []
    for X, kind, N in [("orth", "clements", 40), ("orth", "clements", 200), ("orth", "clements", 700),
                       ("mzi", "butterfly", 128), ("mzi", "butterfly", 1024), ("mzi", "clements", 600)]:
        if kind == "clements":
            g = O.clements(N)
            kw = {}
        else:
            g = O.Geometry(N, int(np.log2(N)), strides=np.arange(int(np.log2(N))))
            kw = dict(strides=g.strides)
        n_p = O.N_PHASE[X]
        p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
        meshes.append(MeshNetworkGPU(N, g.L, g.lens, g.shifts, p, np.zeros(0), X, is_phase=False, **kw))
        geoms.append((g, p))
    server = MeshServer(meshes, hbm_budget=3 << 30, slots=3, direct=False)
    reqs = poisson_requests(meshes, rate=50000.0, n=60, seed=9, b_max=200)
    done, st = server.serve(reqs)
    assert st["requests"] == 60
    seq, _ = server.serve_sequential(_clone(reqs))
    checked = 0
    for r, q in zip(done, seq):
        m = meshes[r.mesh_id]
        assert r.y.dtype == m.dtype
        if r.op == "grad":
            job = GroupJob(m, "grad", x=r.x, target=r.target, tables=server.tables[r.mesh_id], build_tables=False)
        else:
            job = GroupJob(m, "fwd", x=r.x, tables=server.tables[r.mesh_id], build_tables=False)
        MeshGroup([job]).run()
        torch.cuda.synchronize()
        assert torch.equal(job.y, r.y), (r.rid, m.X, m.N, r.op)
        assert O.rel_l2(q.y.cpu().numpy(), r.y.cpu().numpy()) < TOL
        if r.op == "grad":
            assert torch.equal(job.dp, r.dp), (r.rid, m.X, m.N)
            assert O.rel_l2(q.dp.cpu().numpy(), r.dp.cpu().numpy()) < TOL
        if m.X == "orth" and m.N <= 200:
            g, p = geoms[r.mesh_id]
            p64 = p.astype(np.float64)
            x = r.x.cpu().numpy().astype(np.float64)
            y = O.mesh_fwd(g, "orth", p64, None, x)
            assert O.rel_l2(r.y.cpu().numpy(), y) < TOL
            if r.op == "grad":
                yg = r.y.cpu().numpy().astype(np.float64)
                rdp, _, _ = O.mesh_vjp(g, "orth", p64, None, yg, 2 * (yg - r.target.cpu().numpy().astype(np.float64)))
                assert O.rel_l2(r.dp.cpu().numpy().reshape(rdp.shape), rdp) < TOL
            checked += 1
    assert checked >= 5


def test_admission_respects_budget(gpu):
    from meshes_amd.serving import MeshServer, poisson_requests, request_footprint
    meshes = _meshes(seed=4)
    reqs = poisson_requests(meshes, rate=50000.0, n=40, seed=6, b_max=256)
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    need = max(request_footprint(meshes[r.mesh_id], r.op, r.B, ncu) for r in reqs)
    server = MeshServer(meshes, hbm_budget=2 * need, slots=2)
    done, st = server.serve(reqs)
    assert all(r.t_done >= r.t_arrive for r in done)
    assert st["peak_inflight_bytes"] <= 2 * need
    assert server.inflight_bytes == 0


def test_oversized_request_is_rejected(gpu):
    from meshes_amd.serving import MeshServer, poisson_requests
    meshes = _meshes(seed=5)[-2:]
    server = MeshServer(meshes, hbm_budget=1 << 20, slots=1, graphs=False)
    reqs = poisson_requests(meshes, rate=1000.0, n=3, seed=7, b_median=512, b_sigma=0.0)
    with pytest.raises(ValueError):
        server.serve(reqs)


def test_served_equals_sequential_with_splitter_errors(gpu):
    """Meshes with splitter imperfections: the sequential baseline carries p_splitter (and its
    leading dimension) exactly as the served entries do, so both compute the same mesh; both
    agree with the FP64 oracle."""
    from meshes_amd import MeshNetworkGPU, geometry
    from meshes_amd.serving import MeshServer, poisson_requests
    rng = np.random.RandomState(17)
    meshes = []
    for N in (16, 48, 96):
        lens, shifts = geometry.clements(N)
        s = (0.05 * rng.standard_normal((N, N // 2, 2))).astype(np.float32)
        meshes.append(MeshNetworkGPU(N, N, lens, shifts, geometry.uniform_phases(rng, N, N), s, "mzi",
                                     is_phase=False))
    server = MeshServer(meshes, hbm_budget=1 << 30, slots=2)
    reqs = poisson_requests(meshes, rate=20000.0, n=24, seed=6, b_max=40)
    done, _ = server.serve(reqs)
    seq, _ = server.serve_sequential(_clone(reqs))
    for r, q in zip(done, seq):
        assert O.rel_l2(q.y.cpu().numpy(), r.y.cpu().numpy()) < TOL, (r.rid, r.op)
        if r.op == "grad":
            assert O.rel_l2(q.dp.cpu().numpy(), r.dp.cpu().numpy()) < TOL, r.rid
        m = meshes[r.mesh_id]
        g = O.clements(m.N)
        p = m.p_phase.cpu().numpy().astype(np.float64).reshape(m.N, m.N // 2, 2)
        s = m.p_splitter.cpu().numpy().astype(np.float64).reshape(m.N, m.N // 2, 2)
        y = O.mesh_fwd(g, "mzi", p, s, r.x.cpu().numpy().astype(np.complex128))
        assert O.rel_l2(q.y.cpu().numpy(), y) < TOL, (r.rid, m.N)


def test_lone_requests_run_direct(gpu):
    """At low load a request that finds the other slots idle runs as one plain meshprop_run call
    (no plan upload, no grouped launch); its results equal the grouped launch of the same
    request and the FP64 oracle within 1e-5."""
    from meshes_amd.batched import GroupJob, MeshGroup
    from meshes_amd.serving import MeshServer, poisson_requests
    meshes = _meshes(seed=8)[:6]
    server = MeshServer(meshes, hbm_budget=1 << 30, slots=2)
    reqs = poisson_requests(meshes, rate=300.0, n=16, seed=9, b_max=64)
    done, st = server.serve(reqs)
    assert st["direct_launches"] >= 8, st
    for r in done:
        m = meshes[r.mesh_id]
        if r.op == "grad":
            job = GroupJob(m, "grad", x=r.x, target=r.target, tables=server.tables[r.mesh_id], build_tables=False)
        else:
            job = GroupJob(m, "fwd", x=r.x, tables=server.tables[r.mesh_id], build_tables=False)
        MeshGroup([job]).run()
        torch.cuda.synchronize()
        assert O.rel_l2(job.y.cpu().numpy(), r.y.cpu().numpy()) < TOL, (r.rid, m.N, r.op)
        if r.op == "grad":
            assert O.rel_l2(job.dp.cpu().numpy(), r.dp.cpu().numpy()) < TOL, (r.rid, m.N)
        g = O.clements(m.N)
        p = m.p_phase.cpu().numpy().astype(np.float64).reshape(m.N, m.N // 2, 2)
        y = O.mesh_fwd(g, "mzi", p, None, r.x.cpu().numpy().astype(np.complex128))
        assert O.rel_l2(r.y.cpu().numpy(), y) < TOL, (r.rid, m.N)
