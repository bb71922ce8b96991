"""The measured tile table of the plain path (csrc/tile_table.inc, consulted by resolve_tile when a
call brings no tile; the reference's counterpart is gpu/__init__.py:26-35): every distinct tile the
table picks, run at a size that selects it, equals the FP64 oracle within 1e-5 relative L2; and
the sweep that writes the table runs end to end on a small subset.
"""
import os
import re

import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
INC = os.path.join(os.path.dirname(__file__), "..", "meshes_amd", "csrc", "tile_table.inc")
OPS = {0: "fwdprop", 1: "fwddiff", 2: "backdiff", 3: "fwdgrad"}
XS = {0: "mzi", 1: "sym", 2: "orth", 3: "mzi_generic"}


def _entries():
    rows = []
    with open(INC) as f:
        for line in f:
            m = re.match(r"\{(-?\d+), (-?\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\},", line.strip())
            if m:
                rows.append(tuple(int(x) for x in m.groups()))
    return rows


def _cases():
    """One case per distinct (op, X, W, V, waves) of the table: its smallest N bucket, N two below
    the bucket (padded width), a fat (B > 2N) or square-like batch, six layers."""
    seen, out = set(), []
    for op, X, fat, n0, W, V, waves in sorted(_entries(), key=lambda e: e[3]):
        key = (op, X, W, V, waves)
        if key in seen or op not in OPS or X not in XS:
            continue
        if X == 3 and op == 3:
            continue  # fused fwd+grad of lossy crossings needs the scaled-unitary assertion
        seen.add(key)
        N = n0 - 2 if n0 > 8 else n0
        B = 2 * N + 3 if fat else max(2, N // 2 + 1)
        out.append((OPS[op], XS[X], N, B, (W, V, waves)))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c[:4])) + f"-{c[4]}" for c in CASES])
def test_table_tile_parity(gpu, case):
    import meshes_amd._lib as Lb
    from meshes_amd import MeshNetworkGPU
    opname, X, N, B, tile = case
    op = {"fwdprop": Lb.OP_FWDPROP, "fwddiff": Lb.OP_FWDDIFF, "backdiff": Lb.OP_BACKDIFF, "fwdgrad": Lb.OP_FWDGRAD}[opname]
    want = tile
    if tile[0] == 16 and X in ("mzi", "sym") and op in (Lb.OP_BACKDIFF, Lb.OP_FWDGRAD):
        # the NN-specialised W = 16 adjoint picks its workgroup size by batch
        want = (16, tile[1], 8 if B >= 4 * torch.cuda.get_device_properties(0).multi_processor_count else 4)
    assert Lb.resolve_tile(op, X, N, B)[:3] == want  # the call below runs this tile
    rng = np.random.RandomState(N * 7 + B)
    g = O.random_geometry_nn(rng, N, 6)
    n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p)).astype(np.float32)
    s = (0.05 * rng.standard_normal((g.L, N // 2, n_s))).astype(np.float32) if n_s else None
    if X == "mzi_generic":
        s[..., 10:] = 0.0
    real = O.REAL_FIELD[X]
    u = O.unit_cn(rng, B, N, real=real).astype(np.float32 if real else np.complex64)
    mesh = MeshNetworkGPU(N, g.L, g.lens, g.shifts, p, s if s is not None else np.zeros(0), X, is_phase=False)
    p64, s64, u128 = p.astype(np.float64), None if s is None else s.astype(np.float64), u.astype(np.complex128)
    ut = torch.as_tensor(u.T)
    if opname == "fwdprop":
        out = mesh.dot(ut).detach().cpu().numpy().T
        assert O.rel_l2(out, O.mesh_fwd(g, X, p64, s64, u128)) < TOL
    elif opname == "fwddiff":
        dp = rng.standard_normal(p.shape).astype(np.float32)
        du = rng.standard_normal(u.shape) + (0 if real else 1j * rng.standard_normal(u.shape))
        du = du.astype(u.dtype)
        out, dout = mesh.dot(ut, dp=dp, dv=torch.as_tensor(du.T))
        ref, dref = O.mesh_jvp(g, X, p64, s64, u128, dp.astype(np.float64), du.astype(np.complex128))
        assert O.rel_l2(out.detach().cpu().numpy().T, ref) < TOL
        assert O.rel_l2(dout.detach().cpu().numpy().T, dref) < TOL
    elif opname == "backdiff":
        uo = mesh.dot(ut).detach().cpu().numpy().T
        y = O.unit_cn(rng, B, N, real=real).astype(u.dtype)
        gout = (2 * (uo - y)).astype(u.dtype)
        dp, gin = mesh.grad_phi(torch.as_tensor(uo.T), torch.as_tensor(gout.T), vpos="out")
        rdp, rgin, _ = O.mesh_vjp(g, X, p64, s64, uo.astype(np.complex128), gout.astype(np.complex128),
                                  u_in=u128 if X == "mzi_generic" else None)
        assert O.rel_l2(dp.detach().cpu().numpy().reshape(rdp.shape), rdp) < TOL
        assert O.rel_l2(gin.detach().cpu().numpy().T, rgin) < TOL
    else:
        y = O.unit_cn(rng, B, N, real=real).astype(u.dtype)
        J, dp, uo = mesh.grad_target_l2(ut, torch.as_tensor(y.T))
        out = O.mesh_fwd(g, X, p64, s64, u128)
        rdp, _, _ = O.mesh_vjp(g, X, p64, s64, out, 2 * (out - y.astype(np.complex128)), u_in=u128)
        assert O.rel_l2(uo.detach().cpu().numpy().T, out) < TOL
        assert O.rel_l2(dp.detach().cpu().numpy().reshape(rdp.shape), rdp) < TOL


def test_sweep_runs(gpu, tmp_path):
    """The table's writer on a small subset: one line per case, a parseable initializer list."""
    from meshes_amd import tune_tiles
    inc, log = tmp_path / "t.inc", tmp_path / "t.md"
    tune_tiles.main(["--inc", str(inc), "--log", str(log), "--ops", "fwdprop,backdiff", "--xs", "mzi",
                     "--buckets", "8,64"])
    lines = [ln for ln in inc.read_text().splitlines() if ln.startswith("{")]
    assert len(lines) == 8  # 2 ops x 2 shapes x 2 buckets
    assert log.read_text().count("\n| ") >= 8
