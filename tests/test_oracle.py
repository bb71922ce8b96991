"""Pins the FP64 oracle (CPU only).

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong sign,
a transposed operand, a wrong pairing) fails at least one of them:
  * golden vectors written by the compiled reference engine c/dot.cpp (modes 0/1/2);
  * closed forms, unitarity/|det| invariants and the ideal cross/bar special cases of T;
  * central finite differences of T, of the propagated field and of a real loss;
  * dense block-diagonal brute force (gpu/test.py:73-80 construction) for NN and FFT layers;
  * the adjoint identity <g, J du> = Re <J^H g, du>.
"""
import glob
import os

import numpy as np
import pytest

from oracle import meshprop_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ALL_GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
GOLDEN = [p for p in ALL_GOLDEN if "screen" not in os.path.basename(p)]   # slot layout, no screen
PACKED = [p for p in ALL_GOLDEN if "screen" in os.path.basename(p)]   # packed, perms, phase screen


def _geom(d):
    return O.Geometry(int(d["N"]), int(d["L"]), lens=d["lens"], shifts=d["shifts"])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_reference_engine(path):
    d = np.load(path)
    g = _geom(d)
    X = str(d["X"])
    # T of the fixture is built by the oracle's crossing model; check it is what the file holds.
    s = d["s"] if d["s"].size else None
    assert O.rel_l2(O.slot_T(g, X, d["p"], s), d["T"]) < 1e-14
    assert O.rel_l2(O.fwd(g, d["T"], d["u_in"]), d["fwd_out"]) < 1e-13
    u, du = O.jvp(g, d["T"], d["dT"], d["u_in"])
    assert O.rel_l2(u, d["jvp_out"]) < 1e-13
    assert O.rel_l2(du, d["jvp_dout"]) < 1e-13
    rec, gin, dJdT = O.vjp_inverse(g, d["T"], d["fwd_out"], d["g_out"])
    assert O.rel_l2(rec, d["vjp_u_in"]) < 1e-13
    assert O.rel_l2(gin, d["vjp_g_in"]) < 1e-13
    assert O.rel_l2(dJdT[O.active_mask(g)], d["vjp_dJdT"][O.active_mask(g)]) < 1e-13


def test_c1_explicit_unitary():
    """Config C1: the forward pass equals the explicit 16x16 unitary applied to the batch."""
    d = np.load(os.path.join(HERE, "golden", "c1_clements16.npz"))
    g = _geom(d)
    U = O.dense_matrix(g, d["T"])
    assert np.allclose(U.conj().T @ U, np.eye(16), atol=1e-13)
    assert O.rel_l2(d["u_in"] @ U.T, d["fwd_out"]) < 1e-13


@pytest.mark.parametrize("X", ["mzi", "sym", "orth", "mzi_generic"])
def test_crossing_unitarity(X):
    rng = np.random.RandomState(11)
    p = rng.uniform(-7, 7, size=(50, O.N_PHASE[X]))
    s = rng.uniform(-1, 1, size=(50, O.N_SPLIT[X]))
    if X == "mzi_generic":
        s[:, 10:] = 0.0  # unit gains: lossless, so unitary
    T = O.crossing_T(X, p, s if O.N_SPLIT[X] else None)
    I = np.einsum("nij,nkj->nik", T, T.conj())
    assert np.allclose(I, np.eye(2), atol=1e-13)
    assert np.allclose(np.abs(np.linalg.det(T)), 1.0, atol=1e-13)


def test_mzi_closed_form_special_cases():
    # Ideal MZI: T = i e^{i th/2} [[e^{i phi} S, C], [e^{i phi} C, -S]]  (Appendix form of crossing.py:216-221)
    th, ph = 0.7, -1.3
    T = O.crossing_T("mzi", np.array([th, ph]))
    S, C = np.sin(th / 2), np.cos(th / 2)
    ref = 1j * np.exp(1j * th / 2) * np.array([[np.exp(1j * ph) * S, C], [np.exp(1j * ph) * C, -S]])
    assert np.allclose(T, ref, atol=1e-15)
    cross = O.crossing_T("mzi", np.array([0.0, 0.3]))
    assert abs(cross[0, 0]) < 1e-15 and abs(abs(cross[0, 1]) - 1) < 1e-15
    bar = O.crossing_T("mzi", np.array([np.pi, 0.3]))
    assert abs(abs(bar[0, 0]) - 1) < 1e-15 and abs(bar[0, 1]) < 1e-15
    # |T00|^2 = S^2 Cm^2 + C^2 Sp^2 with splitter errors
    a, b = 0.05, -0.03
    T = O.crossing_T("mzi", np.array([th, ph]), np.array([a, b]))
    assert np.isclose(abs(T[0, 0]) ** 2, (S * np.cos(a - b)) ** 2 + (C * np.sin(a + b)) ** 2, atol=1e-15)


def test_generic_matches_mzi_when_lossless():
    """The generic lossy MZI with zero phase errors and unit gains is the MZI with splitters (alpha, beta)."""
    rng = np.random.RandomState(3)
    p = rng.uniform(0, 2 * np.pi, size=(20, 2))
    ab = 0.1 * rng.standard_normal((20, 2))
    s = np.zeros((20, 14))
    s[:, :2] = ab
    assert np.allclose(O.crossing_T("mzi_generic", p, s), O.crossing_T("mzi", p, ab), atol=1e-14)


def test_sym_and_orth_closed_forms():
    th, ph, al = 1.1, 0.4, 0.2
    T = O.crossing_T("sym", np.array([th, ph]), np.array([al]))
    S, C = np.sin(th / 2), np.cos(th / 2)
    # GPU packing T = [[a+ib, ic], [ic, a-ib]] (gpu/crossing.cu:104-106)
    a = S * np.cos(ph) - C * np.sin(ph) * np.sin(2 * al)
    b = S * np.sin(ph) + C * np.cos(ph) * np.sin(2 * al)
    c = C * np.cos(2 * al)
    assert np.allclose(T, [[a + 1j * b, 1j * c], [1j * c, a - 1j * b]], atol=1e-15)
    T = O.crossing_T("orth", np.array([th]))
    assert np.allclose(T, [[S, -C], [C, S]], atol=1e-15)


@pytest.mark.parametrize("X", ["mzi", "sym", "orth", "mzi_generic"])
def test_dT_finite_difference(X):
    rng = np.random.RandomState(5)
    p = rng.uniform(-4, 4, size=(30, O.N_PHASE[X]))
    s = 0.3 * rng.standard_normal((30, O.N_SPLIT[X])) if O.N_SPLIT[X] else None
    dT = O.crossing_T(X, p, s)  # noqa: F841 (shape check below)
    dT = O.crossing_dT(X, p, s)
    h = 1e-6
    for k in range(O.N_PHASE[X]):
        e = np.zeros(O.N_PHASE[X])
        e[k] = h
        fd = (O.crossing_T(X, p + e, s) - O.crossing_T(X, p - e, s)) / (2 * h)
        assert np.allclose(dT[:, k], fd, atol=1e-8), (X, k)


def _fft_dense_layer(N, st, T_l):
    """Block-diagonal FFT layer built as in gpu/test.py:73-77."""
    S = 1 << st
    mats = T_l.transpose(1, 2, 0).reshape(2, 2, N // 2 // S, S).transpose(2, 0, 1, 3)
    blocks = [np.block([[np.diag(M[0, 0]), np.diag(M[0, 1])], [np.diag(M[1, 0]), np.diag(M[1, 1])]]) for M in mats]
    out = np.zeros((N, N), complex)
    for k, Bk in enumerate(blocks):
        out[k * 2 * S:(k + 1) * 2 * S, k * 2 * S:(k + 1) * 2 * S] = Bk
    return out


def _nn_dense_layer(N, sh, ln, T_l):
    """block_diag(eye(sh), T_k..., eye(rest)) as in gpu/test.py:78-80."""
    out = np.eye(N, dtype=complex)
    for k in range(ln):
        i = sh + 2 * k
        out[i:i + 2, i:i + 2] = T_l[sh // 2 + k]
    return out


@pytest.mark.parametrize("kind", ["nn", "fft"])
def test_dense_brute_force(kind):
    rng = np.random.RandomState(7 if kind == "nn" else 8)
    N, L = 32, 9
    if kind == "nn":
        g = O.random_geometry_nn(rng, N, L)
    else:
        g = O.Geometry(N, L, strides=rng.randint(0, 5, size=L))
    T = O.slot_T(g, "mzi", rng.standard_normal((L, N // 2, 2)), 0.1 * rng.standard_normal((L, N // 2, 2)))
    U = np.eye(N, dtype=complex)
    for l in range(L):
        M = (_fft_dense_layer(N, int(g.strides[l]), T[l]) if kind == "fft"
             else _nn_dense_layer(N, int(g.shifts[l]), int(g.lens[l]), T[l]))
        U = M @ U
    assert np.allclose(U, O.dense_matrix(g, T), atol=1e-13)
    u = rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))
    assert O.rel_l2(O.fwd(g, T, u), u @ U.T) < 1e-13


def _loss(g, X, p, s, u, y):
    out = O.mesh_fwd(g, X, p, s, u)
    return np.sum(np.abs(out - y) ** 2)


@pytest.mark.parametrize("X,kind", [("mzi", "nn"), ("sym", "nn"), ("orth", "nn"), ("mzi", "fft"),
                                    ("mzi_generic", "nn")])
def test_jvp_and_vjp_finite_difference(X, kind):
    rng = np.random.RandomState(13)
    N, L, B = 16, 7, 4
    g = O.random_geometry_nn(rng, N, L) if kind == "nn" else O.Geometry(N, L, strides=rng.randint(0, 4, size=L))
    p = rng.standard_normal((L, N // 2, O.N_PHASE[X]))
    s = 0.2 * rng.standard_normal((L, N // 2, O.N_SPLIT[X])) if O.N_SPLIT[X] else None
    if X == "mzi_generic":
        s[..., 10:] = -0.05 + 0.02 * rng.standard_normal((L, N // 2, 4))  # lossy arms
    real = O.REAL_FIELD[X]
    u = rng.standard_normal((B, N)) + (0 if real else 1j * rng.standard_normal((B, N)))
    y = rng.standard_normal((B, N)) + (0 if real else 1j * rng.standard_normal((B, N)))
    dp = rng.standard_normal(p.shape)
    h = 1e-6
    # JVP against FD of the forward map
    out, dout = O.mesh_jvp(g, X, p, s, u, dp)
    fd = (O.mesh_fwd(g, X, p + h * dp, s, u) - O.mesh_fwd(g, X, p - h * dp, s, u)) / (2 * h)
    assert O.rel_l2(dout, fd) < 1e-7
    # VJP: directional derivative of J = sum |Uu - y|^2 along dp and along du
    gout = 2 * (out - y)
    lossless = X != "mzi_generic"
    dJdp, gin, _ = O.mesh_vjp(g, X, p, s, out, gout, u_in=None if lossless else u)
    fdJ = (_loss(g, X, p + h * dp, s, u, y) - _loss(g, X, p - h * dp, s, u, y)) / (2 * h)
    assert np.isclose(np.sum(dJdp * dp), fdJ, rtol=1e-6, atol=1e-9)
    du = rng.standard_normal(u.shape) + (0 if real else 1j * rng.standard_normal(u.shape))
    fdu = (_loss(g, X, p, s, u + h * du, y) - _loss(g, X, p, s, u - h * du, y)) / (2 * h)
    assert np.isclose(np.real(np.sum(np.conj(gin) * du)), fdu, rtol=1e-6, atol=1e-9)


def test_stored_equals_inverse_on_unitary():
    rng = np.random.RandomState(17)
    g = O.clements(12)
    p = rng.uniform(0, 2 * np.pi, (12, 6, 2))
    u = O.unit_cn(rng, 6, 12)
    out = O.mesh_fwd(g, "mzi", p, None, u)
    gout = rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)
    a = O.mesh_vjp(g, "mzi", p, None, out, gout)
    b = O.mesh_vjp(g, "mzi", p, None, out, gout, u_in=u)
    assert O.rel_l2(a[0], b[0]) < 1e-12 and O.rel_l2(a[1], b[1]) < 1e-12 and O.rel_l2(a[2], u) < 1e-12


def test_planted_mistake_is_caught():
    """A sign slip in dT/dtheta (the kind of error the FD pin exists for) fails the FD check."""
    rng = np.random.RandomState(19)
    p = rng.uniform(-3, 3, (10, 2))
    s = 0.2 * rng.standard_normal((10, 2))
    h = 1e-6
    good = O.crossing_dT("mzi", p, s)
    fd = (O.crossing_T("mzi", p + [h, 0], s) - O.crossing_T("mzi", p - [h, 0], s)) / (2 * h)
    planted = good.copy()
    planted[:, 0, 0, 1] *= -1
    assert np.allclose(good[:, 0], fd, atol=1e-8)
    assert not np.allclose(planted[:, 0], fd, atol=1e-8)


def _packed_setup(d):
    N = int(d["N"])
    lens, shifts = [int(x) for x in d["lens"]], [int(x) for x in d["shifts"]]
    perms = d["perms"] if d["perms"].size else None
    X = str(d["X"])
    T = O.crossing_T(X, d["p"]).transpose(1, 2, 0)
    dT = np.einsum("nkij,nk->ijn", O.crossing_dT(X, d["p"]), d["dp"])
    return N, lens, shifts, perms, X, T, dT, int(d["pos"])


@pytest.mark.parametrize("path", PACKED, ids=[os.path.basename(p)[:-4] for p in PACKED])
def test_golden_packed_with_perms_and_screen(path):
    """structured_run (permutations, phase screen at either end) against the compiled
    reference engine c/dot.cpp, modes 0/1/2, and the dense brute-force product."""
    d = np.load(path)
    N, lens, shifts, perms, X, T, dT, pos = _packed_setup(d)
    out = O.structured_run(N, lens, shifts, T, d["u_in"], 0, ph=d["ph"], perms=perms, pos=pos)
    assert O.rel_l2(out, d["fwd_out"]) < 1e-13
    U = O.structured_dense(N, lens, shifts, T, perms=perms, ph=d["ph"], pos=pos)
    assert O.rel_l2(d["u_in"] @ U.T, d["fwd_out"]) < 1e-13
    o1, do1 = O.structured_run(N, lens, shifts, T, d["u_in"], 1, dT=dT, ph=d["ph"], dph=d["dph"], perms=perms,
                               pos=pos)
    assert O.rel_l2(o1, d["jvp_out"]) < 1e-13 and O.rel_l2(do1, d["jvp_dout"]) < 1e-13
    pbk = None if perms is None else np.array([np.argsort(q) for q in perms])
    rec, gin, dJdT, dph = O.structured_run(N, lens, shifts, T, d["fwd_out"], 2, dv=d["g_out"], ph=d["ph"], perms=pbk,
                                           pos=pos)
    assert O.rel_l2(rec, d["vjp_u_in"]) < 1e-13 and O.rel_l2(gin, d["vjp_g_in"]) < 1e-13
    assert O.rel_l2(dJdT, d["vjp_dJdT"]) < 1e-13 and O.rel_l2(dph, d["vjp_dph"]) < 1e-13


@pytest.mark.parametrize("pos", [0, 1])
def test_phase_screen_finite_difference(pos):
    """dJ/dphi of the packed adjoint and the screen tangent against central differences."""
    rng = np.random.RandomState(31 + pos)
    N, B = 8, 5
    g = O.clements(N)
    lens, shifts = list(g.lens), list(g.shifts)
    p = rng.uniform(0, 2 * np.pi, (sum(lens), 2))
    T = O.crossing_T("mzi", p).transpose(1, 2, 0)
    ph = rng.uniform(0, 2 * np.pi, N)
    u = O.unit_cn(rng, B, N)
    y = O.unit_cn(rng, B, N)
    dph = rng.standard_normal(N)
    h = 1e-6

    def loss(phv):
        return np.sum(np.abs(O.structured_run(N, lens, shifts, T, u, 0, ph=phv, pos=pos) - y) ** 2)
    out, dout = O.structured_run(N, lens, shifts, T, u, 1, dT=np.zeros_like(T), ph=ph, dph=dph, pos=pos)
    fd = (O.structured_run(N, lens, shifts, T, u, 0, ph=ph + h * dph, pos=pos)
          - O.structured_run(N, lens, shifts, T, u, 0, ph=ph - h * dph, pos=pos)) / (2 * h)
    assert O.rel_l2(dout, fd) < 1e-7
    _, _, _, dJ = O.structured_run(N, lens, shifts, T, out, 2, dv=2 * (out - y), ph=ph, pos=pos)
    fdJ = (loss(ph + h * dph) - loss(ph - h * dph)) / (2 * h)
    assert np.isclose(np.sum(dJ * dph), fdJ, rtol=1e-6, atol=1e-9)


def test_butterfly_perms_pair_stride_neighbours():
    """After the gathers of butterfly_perms, layer i couples waveguides 2^tz(i+1) apart for odd i
    (butterfly.py:46-51) and adjacent ones for even i; the last gather restores the order."""
    N = 32
    lens, shifts, perms = O.butterfly_perms(N)
    pos = np.arange(N)
    for i in range(N - 1):
        pos = pos[perms[i]]
        a, b = pos[0::2], pos[1::2]
        s = 1 if i % 2 == 0 else ((i + 1) & -(i + 1))
        assert np.all(b - a == s) and np.all((a & s) == 0)
    assert np.array_equal(pos[perms[N - 1]], np.arange(N))


@pytest.mark.parametrize("X", ["mzi", "sym", "orth", "mzi_generic"])
def test_d2T_finite_difference(X):
    rng = np.random.RandomState(41)
    n_p = O.N_PHASE[X]
    p = rng.uniform(0, 2 * np.pi, (5, n_p))
    s = 0.1 * rng.standard_normal((5, O.N_SPLIT[X])) if O.N_SPLIT[X] else None
    d2 = O.crossing_d2T(X, p, s)
    h = 1e-6
    for m in range(n_p):
        e = np.zeros(n_p)
        e[m] = h
        fd = (O.crossing_dT(X, p + e, s) - O.crossing_dT(X, p - e, s)) / (2 * h)   # [5, k, 2, 2]
        assert np.allclose(d2[:, :, m], fd, atol=1e-8), (X, m)
    assert np.allclose(d2, np.swapaxes(d2, 1, 2))


@pytest.mark.parametrize("X,kind", [("mzi", "nn"), ("sym", "nn"), ("mzi", "fft"), ("orth", "nn")])
def test_hvp_finite_difference_of_gradient(X, kind):
    """hvp_inverse against central differences (h = 1e-5) of the FP64 gradient along q (real
    fields for orth)."""
    rng = np.random.RandomState(43)
    N, L, B = 12, 9, 5
    g = O.random_geometry_nn(rng, N, L) if kind == "nn" else O.Geometry(16, 6, strides=rng.randint(0, 4, size=6))
    N = g.N
    n_p = O.N_PHASE[X]
    p = rng.uniform(0, 2 * np.pi, (g.L, N // 2, n_p))
    s = 0.1 * rng.standard_normal((g.L, N // 2, O.N_SPLIT[X]))
    u = O.unit_cn(rng, B, N, real=O.REAL_FIELD[X])
    t = O.unit_cn(rng, B, N, real=O.REAL_FIELD[X])
    q = rng.standard_normal(p.shape)

    def grad(pp):
        out = O.mesh_fwd(g, X, pp, s, u)
        return O.mesh_vjp(g, X, pp, s, out, 2 * (out - t))[0]
    gr, hv = O.hvp_inverse(g, X, p, s, u, q, t)
    assert O.rel_l2(gr, grad(p)) < 1e-12
    h = 1e-5
    fd = (grad(p + h * q) - grad(p - h * q)) / (2 * h)
    assert O.rel_l2(hv, fd) < 1e-7


def _orth_reference_walk(g, p, u_in, du_in, dp, g_out):
    """The reference GPU kernel's own orth formulas, restated on one NN mesh:
    forward matmult_orth u1' = T0 u1 + T32 u2, u2' = -T32 u1 + T0 u2 with T0 = S, T32 = -C
    (gpu/crossing.cu:132-139, 230-236); tangent matmult_d_orth with dth = dp/2
    (gpu/crossing.cu:142-147, 264-272); adjoint matmult_bk_orth: reconstruct
    x1 = T0 u1 - T32 u2, x2 = T32 u1 + T0 u2, then dth += T0 (x2 g1 - x1 g2) - T32 (x1 g1 + x2 g2)
    with g the layer-output adjoint, g <- (T0 g1 - T32 g2, T32 g1 + T0 g2) (gpu/crossing.cu:320-333),
    and dp = dth / 2 (dp_mzi_orth, gpu/crossing.cu:200-203).  Uses no function of the oracle."""
    u, du = u_in.copy(), du_in.copy()
    for l in range(g.L):
        m, i, j = g.pairs(l)
        th = p[l, m, 0]
        T0, T32, dth = np.sin(th / 2), -np.cos(th / 2), 0.5 * dp[l, m, 0]
        u1, u2, d1, d2 = u[:, i].copy(), u[:, j].copy(), du[:, i].copy(), du[:, j].copy()
        du[:, i] = T0 * d1 + T32 * d2 + dth * (-T32 * u1 + T0 * u2)
        du[:, j] = -T32 * d1 + T0 * d2 + dth * (-T0 * u1 - T32 * u2)
        u[:, i], u[:, j] = T0 * u1 + T32 * u2, -T32 * u1 + T0 * u2
    out, dout = u.copy(), du.copy()
    grad = np.zeros_like(p)
    gg = g_out.copy()
    for l in range(g.L - 1, -1, -1):
        m, i, j = g.pairs(l)
        th = p[l, m, 0]
        T0, T32 = np.sin(th / 2), -np.cos(th / 2)
        u1, u2, g1, g2 = u[:, i].copy(), u[:, j].copy(), gg[:, i].copy(), gg[:, j].copy()
        x1, x2 = T0 * u1 - T32 * u2, T32 * u1 + T0 * u2
        d_th = np.sum(T0 * (x2 * g1 - x1 * g2) - T32 * (x1 * g1 + x2 * g2), axis=0)
        grad[l, m, 0] = 0.5 * d_th
        u[:, i], u[:, j] = x1, x2
        gg[:, i], gg[:, j] = T0 * g1 - T32 * g2, T32 * g1 + T0 * g2
    return out, dout, grad, gg


def _orth_case():
    rng = np.random.RandomState(41)
    g = O.random_geometry_nn(rng, 14, 9)
    p = rng.uniform(0, 2 * np.pi, (g.L, 7, 1))
    u = rng.standard_normal((5, 14))
    du = rng.standard_normal((5, 14))
    dp = rng.standard_normal(p.shape)
    gout = rng.standard_normal((5, 14))
    return g, p, u, du, dp, gout


def test_orth_matches_reference_kernel_formulas():
    """The oracle's orth forward, JVP and gradient equal the reference kernel's own orth
    formulas (an independent statement of the orth convention: T, dth = dp/2, and the adjoint
    accumulation of gpu/crossing.cu:320-333)."""
    g, p, u, du, dp, gout = _orth_case()
    out, dout, grad, gin = _orth_reference_walk(g, p, u, du, dp, gout)
    o_out, o_dout = O.mesh_jvp(g, "orth", p, None, u, dp, du)
    assert np.max(np.abs(o_out - out)) < 1e-12 and np.max(np.abs(o_dout - dout)) < 1e-12
    o_dp, o_gin, o_rec = O.mesh_vjp(g, "orth", p, None, out, gout)
    assert np.max(np.abs(o_dp - grad)) < 1e-12
    assert np.max(np.abs(o_gin - gin)) < 1e-12 and np.max(np.abs(o_rec - u)) < 1e-12


def test_orth_planted_transpose_is_caught(monkeypatch):
    """A transposed orth T ([[S, C], [-C, S]], still orthogonal, so unitarity and the oracle's
    own self-consistent FD checks pass) fails the comparison with the reference formulas."""
    g, p, u, du, dp, gout = _orth_case()
    out, dout, grad, gin = _orth_reference_walk(g, p, u, du, dp, gout)
    good = O.crossing_T

    def transposed(X, pp, s=None):
        T = good(X, pp, s)
        return np.swapaxes(T, -1, -2).copy() if X == "orth" else T
    monkeypatch.setattr(O, "crossing_T", transposed)
    o_dp, _, _ = O.mesh_vjp(g, "orth", p, None, out, gout)
    assert np.max(np.abs(o_dp - grad)) > 1e-3
    o_out = O.mesh_fwd(g, "orth", p, None, u)
    assert np.max(np.abs(o_out - out)) > 1e-3
