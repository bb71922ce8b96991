"""The heterogeneous grouped launch at its own size: the seed-0 mix of 64 meshes (half Clements of
depth N, half butterfly of log2 N stride layers; N uniform over {32, 64, 128, 256, 512}; per-mesh
batch log-uniform in [1, 2048]; MZI phases U[0, 2 pi); fwd+grad of J = sum |U x - t|^2 per mesh),
tiled by the autotuner, captured as a hipGraph and replayed.

Bar: every job's output and gradient within 1e-5 relative L2 of the per-mesh plain launches (one
launch sequence per mesh, the reference's call pattern gpu/mesh.py:158-231) and, for the meshes of
N <= 128, of the FP64 oracle; a mesh run alone gives bitwise the same results as inside the mix.
No reference analog exists for the grouped form itself."""
import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rel(a, b):
    a = a.detach().cpu().numpy().astype(np.complex128)
    b = b.detach().cpu().numpy().astype(np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _oracle_geometry(m):
    if m.fft:
        return O.Geometry(m.N, m.L, strides=m.strides.cpu().numpy())
    return O.Geometry(m.N, m.L, lens=m.lens.cpu().numpy(), shifts=m.shifts.cpu().numpy())


def test_c4_mix_tuned_captured(gpu):
    from meshes_amd import workloads as WL
    from meshes_amd.autotune import tune_group
    from meshes_amd.batched import GroupJob, MeshGroup, PlainRunner
    specs = WL.mix_specs(seed=0)
    assert len(specs) == 64 and {t for t, _, _ in specs} == {"clements", "butterfly"}
    assert max(B for _, _, B in specs) <= 2048 and {N for _, N, _ in specs} <= set(WL.MIX_N)
    jobs = WL.mix_jobs(specs)
    print("c4: tuning the 64-mesh mix", flush=True)
    cfg, best_ms, sweep = tune_group(jobs, reps=1)
    assert len(sweep) > 1 and best_ms > 0 and cfg
    print(f"c4: tuned {cfg} best {best_ms:.3f} ms over {len(sweep)} candidates", flush=True)
    grp = MeshGroup(jobs)
    grp.capture()
    for j in jobs:  # the replay must write every output
        j.y.fill_(float("nan"))
        j.dp.fill_(7.0)
    torch.cuda.synchronize()
    grp.run()
    torch.cuda.synchronize()
    # per-mesh plain launches (library default tiles)
    plain = [GroupJob(j.mesh, "grad", x=j.x, target=j.target) for j in jobs]
    PlainRunner(plain).run()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(jobs, plain)):
        assert _rel(a.y, b.y) < TOL, (i, specs[i])
        assert _rel(a.dp, b.dp) < TOL, (i, specs[i])
    # FP64 oracle on the narrow meshes
    n_checked = 0
    for i, j in enumerate(jobs):
        m = j.mesh
        if m.N > 128:
            continue
        g = _oracle_geometry(m)
        p = m.p_phase.cpu().numpy().astype(np.float64).reshape(m.L, m.N // 2, 2)
        x = j.x.cpu().numpy().astype(np.complex128)
        y = O.mesh_fwd(g, "mzi", p, None, x)
        assert O.rel_l2(j.y.cpu().numpy(), y) < TOL, (i, specs[i])
        yk = j.y.cpu().numpy().astype(np.complex128)
        rdp, _, _ = O.mesh_vjp(g, "mzi", p, None, yk, 2 * (yk - j.target.cpu().numpy()))
        assert O.rel_l2(j.dp.cpu().numpy().reshape(rdp.shape), rdp) < TOL, (i, specs[i])
        n_checked += 1
    assert n_checked >= 20
    # alone == in the mix, bitwise, on one job of every width in the mix (and its tuned tile)
    picked = {}
    for i, (t, N, B) in enumerate(specs):
        picked.setdefault((N, t), i)
    assert len(picked) >= 4
    for key, i in sorted(picked.items()):
        j = jobs[i]
        ja = GroupJob(j.mesh, "grad", x=j.x, target=j.target, tile=j.tile)
        MeshGroup([ja]).run()
        torch.cuda.synchronize()
        assert torch.equal(ja.y, j.y), key
        assert torch.equal(ja.dp, j.dp), key
