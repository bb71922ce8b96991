"""The phase screen (is_phase: exp(i phi) at the mesh input or output, mesh.py:38-41 and
c/dot.cpp:97-143) in the autograd function and the fused fwd+grad call, against the FP64 oracle.

The oracle side composes the pinned mesh oracle (mesh_fwd / mesh_vjp) with the screen applied in
numpy; the screen phases' gradient is also checked against central differences of the oracle's
FP64 loss, which pins it independently of any adjoint formula.  Bar: 1e-5 relative L2.
"""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _problem(X, N, L, B, pos, seed):
    from meshes_amd import MeshNetworkGPU
    rng = np.random.RandomState(seed)
    g = O.random_geometry_nn(rng, N, L)
    n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
    ph = rng.uniform(0, 2 * np.pi, N).astype(np.float32)
    s = None
    if n_s:
        s = (0.05 * rng.standard_normal((g.L, N // 2, n_s))).astype(np.float32)
        if X == "mzi_generic":
            s[..., 10:] = np.float32(np.log(10 ** (-0.1 / 20)))  # balanced 0.1 dB per crossing
    u = O.unit_cn(rng, B, N).astype(np.complex64)
    y = O.unit_cn(rng, B, N).astype(np.complex64)
    mesh = MeshNetworkGPU(N, g.L, g.lens, g.shifts, np.concatenate([p.ravel(), ph]),
                          s if s is not None else np.zeros(0), X, phi_pos=pos, is_phase=True)
    return g, p, ph, s, u, y, mesh


def _oracle_g(g, X, p, ph, s, u, pos, gfun):
    """Output and gradients (mesh phases, screen phases, input) of a loss with dJ/dout = gfun(out)."""
    p64, s64 = p.astype(np.float64), None if s is None else s.astype(np.float64)
    e = np.exp(1j * ph.astype(np.float64))
    u128 = u.astype(np.complex128)
    if pos == "in":
        xs = u128 * e[None, :]
        out = O.mesh_fwd(g, X, p64, s64, xs)
        gout = gfun(out)
        rdp, gxs, _ = O.mesh_vjp(g, X, p64, s64, out, gout, u_in=xs)
        dph = np.sum(np.imag(gxs * np.conj(xs)), axis=0)
        gx = gxs * np.conj(e)[None, :]
    else:
        z = O.mesh_fwd(g, X, p64, s64, u128)
        out = z * e[None, :]
        gout = gfun(out)
        dph = np.sum(np.imag(gout * np.conj(out)), axis=0)
        rdp, gx, _ = O.mesh_vjp(g, X, p64, s64, z, gout * np.conj(e)[None, :], u_in=u128)
    return out, rdp, dph, gx


def _oracle(g, X, p, ph, s, u, y, pos):
    """Output, loss J = sum |out - y|^2 and its gradients (mesh phases, screen phases, input)."""
    y128 = y.astype(np.complex128)
    out, rdp, dph, gx = _oracle_g(g, X, p, ph, s, u, pos, lambda o: 2 * (o - y128))
    return out, float(np.sum(np.abs(out - y128) ** 2)), rdp, dph, gx


def _fd_screen(g, X, p, ph, s, u, y, pos, rng, h=1e-4):
    """Directional central difference of the FP64 loss along a random screen direction."""
    d = rng.standard_normal(ph.shape)
    jp = _oracle(g, X, p, ph.astype(np.float64) + h * d, s, u, y, pos)[1]
    jm = _oracle(g, X, p, ph.astype(np.float64) - h * d, s, u, y, pos)[1]
    return d, (jp - jm) / (2 * h)


@pytest.mark.parametrize("pos", ["in", "out"])
@pytest.mark.parametrize("X", ["mzi", "mzi_generic"])
def test_autograd_screen(gpu, X, pos):
    from meshes_amd.autograd import mesh_apply
    g, p, ph, s, u, y, mesh = _problem(X, 128, 10, 40, pos, seed=3 if pos == "in" else 4)
    pt = torch.nn.Parameter(mesh.p_phase.clone())
    xt = torch.as_tensor(u).cuda().requires_grad_(True)
    out = mesh_apply(mesh, pt, xt, ckpt_every=4 if X == "mzi_generic" else 0)
    loss = torch.sum(torch.abs(out - torch.as_tensor(y).cuda()) ** 2)
    loss.backward()
    rout, rJ, rdp, rdph, rgx = _oracle(g, X, p, ph, s, u, y, pos)
    assert O.rel_l2(out.detach().cpu().numpy(), rout) < TOL
    gp = pt.grad.cpu().numpy()
    n_m = mesh.n_cr * mesh.X_np
    assert O.rel_l2(gp[:n_m].reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(gp[n_m:], rdph) < TOL
    assert O.rel_l2(xt.grad.cpu().numpy(), rgx) < TOL
    d, fd = _fd_screen(g, X, p, ph, s, u, y, pos, np.random.RandomState(11))
    assert abs(np.dot(gp[n_m:].astype(np.float64), d) - fd) < 1e-4 * max(1.0, abs(fd))


@pytest.mark.parametrize("pos", ["in", "out"])
@pytest.mark.parametrize("X", ["mzi", "sym", "mzi_generic"])
def test_fused_fwdgrad_screen(gpu, X, pos):
    g, p, ph, s, u, y, mesh = _problem(X, 256, 12, 33, pos, seed=5 if pos == "in" else 6)
    J, dp, uo = mesh.grad_target_l2(torch.as_tensor(u.T).cuda(), torch.as_tensor(y.T).cuda())
    rout, rJ, rdp, rdph, _ = _oracle(g, X, p, ph, s, u, y, pos)
    assert O.rel_l2(uo.T.cpu().numpy(), rout) < TOL
    assert abs(float(J) - rJ) < 1e-4 * rJ
    gp = dp.cpu().numpy()
    n_m = mesh.n_cr * mesh.X_np
    assert O.rel_l2(gp[:n_m].reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(gp[n_m:], rdph) < TOL
    # accumulating form: a second call adds the same gradient
    acc = dp.clone()
    mesh.grad_target_l2(torch.as_tensor(u.T).cuda(), torch.as_tensor(y.T).cuda(), out_dp=acc)
    assert O.rel_l2(acc.cpu().numpy(), 2 * gp) < TOL


def _check_dp(dp, mesh, rdp, rdph):
    gp = dp.cpu().numpy()
    n_m = mesh.n_cr * mesh.X_np
    assert O.rel_l2(gp[:n_m].reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(gp[n_m:], rdph) < TOL


@pytest.mark.parametrize("pos", ["in", "out"])
def test_grouped_screen(gpu, pos):
    """Grouped jobs on screened meshes (fwd; grad with a target and with a given dJ/dy; backdiff
    with the input reconstructed) beside an unscreened job, against the oracle; the per-mesh plain
    launches agree; captured-graph replays equal the eager launch."""
    from meshes_amd import MeshNetworkGPU, geometry
    from meshes_amd.batched import GroupJob, MeshGroup, run_plain
    probs = [_problem("mzi", N, L, B, pos, seed=20 + k) for k, (N, L, B) in
             enumerate([(64, 9, 21), (128, 12, 40), (96, 7, 33), (256, 10, 17)])]
    c = lambda a: torch.as_tensor(np.ascontiguousarray(a.astype(np.complex64))).cuda()  # noqa: E731
    (g0, p0, ph0, s0, u0, y0, m0), (g1, p1, ph1, s1, u1, y1, m1), (g2, p2, ph2, s2, u2, y2, m2), \
        (g3, p3, ph3, s3, u3, y3, m3) = probs
    rng = np.random.RandomState(77)
    dy2 = (rng.standard_normal(u2.shape) + 1j * rng.standard_normal(u2.shape)).astype(np.complex64)
    dy3 = (rng.standard_normal(u3.shape) + 1j * rng.standard_normal(u3.shape)).astype(np.complex64)
    # backdiff job: the screened mesh's true output of u3
    y3_out = _oracle_g(g3, "mzi", p3, ph3, s3, u3, pos, lambda o: o)[0].astype(np.complex64)
    lens, shifts = geometry.clements(32)
    plain = MeshNetworkGPU(32, 32, lens, shifts, geometry.uniform_phases(rng, 32, 32), np.zeros(0), "mzi",
                           is_phase=False)
    xp = c(O.unit_cn(rng, 50, 32))

    def jobs():
        return [GroupJob(m0, "fwd", x=c(u0)),
                GroupJob(m1, "grad", x=c(u1), target=c(y1), gx=torch.empty_like(c(u1))),
                GroupJob(m2, "grad", x=c(u2), dy=c(dy2), gx=torch.empty_like(c(u2))),
                GroupJob(m3, "backdiff", x=torch.empty_like(c(u3)), y=c(y3_out), dy=c(dy3),
                         gx=torch.empty_like(c(u3))),
                GroupJob(plain, "fwd", x=xp)]

    js = jobs()
    grp = MeshGroup(js)
    grp.run()
    torch.cuda.synchronize()
    assert O.rel_l2(js[0].y.cpu().numpy(), _oracle_g(g0, "mzi", p0, ph0, s0, u0, pos, lambda o: o)[0]) < TOL
    rout, _, rdp, rdph, rgx = _oracle(g1, "mzi", p1, ph1, s1, u1, y1, pos)
    assert O.rel_l2(js[1].y.cpu().numpy(), rout) < TOL
    _check_dp(js[1].dp, m1, rdp, rdph)
    assert O.rel_l2(js[1].gx.cpu().numpy(), rgx) < TOL
    rout, rdp, rdph, rgx = _oracle_g(g2, "mzi", p2, ph2, s2, u2, pos, lambda o: dy2.astype(np.complex128))
    assert O.rel_l2(js[2].y.cpu().numpy(), rout) < TOL
    _check_dp(js[2].dp, m2, rdp, rdph)
    assert O.rel_l2(js[2].gx.cpu().numpy(), rgx) < TOL
    _, rdp, rdph, rgx = _oracle_g(g3, "mzi", p3, ph3, s3, u3, pos, lambda o: dy3.astype(np.complex128))
    _check_dp(js[3].dp, m3, rdp, rdph)
    assert O.rel_l2(js[3].gx.cpu().numpy(), rgx) < TOL
    assert O.rel_l2(js[3].x.cpu().numpy(), u3) < TOL
    # the per-mesh plain launches
    pj = jobs()
    run_plain(pj)
    torch.cuda.synchronize()
    for a, b in zip(js, pj):
        assert O.rel_l2(b.y.cpu().numpy(), a.y.cpu().numpy()) < TOL
        if a.dp is not None:
            assert O.rel_l2(b.dp.cpu().numpy(), a.dp.cpu().numpy()) < TOL
        if a.gx is not None:
            assert O.rel_l2(b.gx.cpu().numpy(), a.gx.cpu().numpy()) < TOL
    # graph replays against the eager results
    want = [(j.y.clone(), None if j.dp is None else j.dp.clone(), None if j.gx is None else j.gx.clone())
            for j in js]
    grp.capture()
    for _ in range(3):
        for j in js:
            if j.op != "backdiff":  # a backdiff job's y is its input
                j.y.zero_()
            if j.gx is not None:
                j.gx.zero_()
            if j.dp is not None:
                j.dp.fill_(7.0)
        grp.run()
        torch.cuda.synchronize()
        for j, (y, dp, gx) in zip(js, want):
            if j.op != "backdiff":
                assert torch.equal(j.y, y)
            if dp is not None:
                assert torch.equal(j.dp, dp)
            if gx is not None:
                assert torch.equal(j.gx, gx)


@pytest.mark.parametrize("pos", ["in", "out"])
def test_served_screen(gpu, pos):
    """The server with screened and unscreened meshes: every request equals the mesh's own
    methods (oracle-checked above) and the sequential baseline."""
    from meshes_amd import MeshNetworkGPU, geometry
    from meshes_amd.serving import MeshServer, Request
    rng = np.random.RandomState(31)
    meshes = [_problem("mzi", N, L, 1, pos, seed=40 + N)[-1] for N, L in [(64, 20), (200, 16)]]
    lens, shifts = geometry.clements(48)
    meshes.append(MeshNetworkGPU(48, 48, lens, shifts, geometry.uniform_phases(rng, 48, 48), np.zeros(0), "mzi",
                                 is_phase=False))
    reqs = []
    for k in range(12):
        mid = k % 3
        N = meshes[mid].N
        B = int(rng.randint(1, 40))
        x = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
        t = torch.as_tensor(O.unit_cn(rng, B, N).astype(np.complex64)).cuda()
        reqs.append(Request(k, mid, "grad" if k % 2 else "fwd", x, t if k % 2 else None, t_arrive=0.0))
    server = MeshServer(meshes, hbm_budget=1 << 30, slots=2)
    done, st = server.serve(reqs)
    seq, _ = server.serve_sequential([Request(r.rid, r.mesh_id, r.op, r.x, r.target, 0.0) for r in done])
    for r, q in zip(done, seq):
        m = meshes[r.mesh_id]
        if r.op == "grad":
            _, dp, uo = m.grad_target_l2(r.x.T, r.target.T)
            assert O.rel_l2(r.y.cpu().numpy(), uo.T.cpu().numpy()) < TOL
            assert O.rel_l2(r.dp.cpu().numpy(), dp.cpu().numpy()) < TOL
            assert O.rel_l2(q.dp.cpu().numpy(), dp.cpu().numpy()) < TOL
        else:
            y = m.dot(r.x.T).T
            assert O.rel_l2(r.y.cpu().numpy(), y.cpu().numpy()) < TOL
        assert O.rel_l2(q.y.cpu().numpy(), r.y.cpu().numpy()) < TOL
