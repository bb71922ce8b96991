"""The six reference-shaped launchers (meshprop_fwdprop/_fwddiff/_backdiff and their _ft
variants) called as a reference maintainer binds them: ctypes argtypes exactly as INTEGRATION.md
section 1 writes them, the reference argument order (gpu/fwdprop.cu:25-27, gpu/fwddiff.cu:24-27,
gpu/backdiff.cu:24-27), backdiff in place (u_in == u_out and dJdu_in == dJdu_out, as
gpu/mesh.py:228-231 calls it) and dp accumulated into a pre-filled buffer (gpu/mesh.py:214).
Bar: 1e-5 relative L2 against the FP64 oracle."""
import ctypes
import os

import numpy as np
import pytest
import torch

from oracle import meshprop_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mp(gpu):
    # -- the binding of INTEGRATION.md section 1, on its own library handle --
    _mp = ctypes.CDLL(os.path.join(ROOT, "meshes_amd", "libmeshprop.so"))
    _i, _p, _z = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
    _mp.meshprop_fwdprop.argtypes = [_i, _i, _i, _i, _p, _p, _p, _i, _p, _i, _p, _p, _i, _i, _p, _z, _p]
    _mp.meshprop_fwddiff.argtypes = [_i, _i, _i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _i, _i, _p, _z, _p]
    _mp.meshprop_backdiff.argtypes = [_i, _i, _i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _i, _i, _p, _z, _p]
    _mp.meshprop_fwdprop_ft.argtypes = [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i, _i, _p, _z, _p]
    _mp.meshprop_fwddiff_ft.argtypes = [_i, _i, _i, _i, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _i, _i, _p, _z, _p]
    _mp.meshprop_backdiff_ft.argtypes = [_i, _i, _i, _i, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _i, _i, _p, _z,
                                         _p]
    _mp.meshprop_workspace_size.argtypes = [_i, _i, _i, _i, _i, _i, _p]
    _mp.meshprop_workspace_size.restype = _z
    _mp.meshprop_status_string.restype = ctypes.c_char_p
    return _mp


X_CODE = {"mzi": 0, "sym": 1, "orth": 2}


def _ptr(t):
    return None if (t is None or t.numel() == 0) else t.data_ptr()


class Mesh:
    """Device buffers of one mesh in the reference's layout: int32 geometry, slot-layout p and s."""

    def __init__(self, X, kind, N, L, B, seed):
        rng = np.random.RandomState(seed)
        self.X, self.fft, self.N, self.B = X, kind == "fft", N, B
        if self.fft:
            self.g = O.Geometry(N, L, strides=rng.randint(0, int(np.log2(N)), size=L))
        else:
            self.g = O.random_geometry_nn(rng, N, L)
        self.L = self.g.L
        n_p, n_s = O.N_PHASE[X], O.N_SPLIT[X]
        self.p = rng.uniform(0, 2 * np.pi, (self.L, N // 2, n_p)).astype(np.float32)
        self.s = (0.05 * rng.standard_normal((self.L, N // 2, n_s))).astype(np.float32) if n_s else None
        real = O.REAL_FIELD[X]
        self.dtype = np.float32 if real else np.complex64
        self.u = O.unit_cn(rng, B, N, real=real).astype(self.dtype)
        self.du = O.unit_cn(rng, B, N, real=real).astype(self.dtype)
        self.dp_dir = rng.standard_normal(self.p.shape).astype(np.float32)
        self.gout = O.unit_cn(rng, B, N, real=real).astype(self.dtype)
        self.dp0 = rng.standard_normal(self.p.size).astype(np.float32)
        dev = torch.device("cuda", 0)
        i32 = dict(dtype=torch.int32, device=dev)
        if self.fft:
            self.strides = torch.as_tensor(np.asarray(self.g.strides), **i32)
        else:
            self.lens = torch.as_tensor(np.asarray(self.g.lens), **i32)
            self.shifts = torch.as_tensor(np.asarray(self.g.shifts), **i32)
        self.p_d = torch.as_tensor(self.p.reshape(-1), device=dev)
        self.s_d = torch.as_tensor(self.s.reshape(-1), device=dev) if self.s is not None else None
        self.ldp, self.lds = N // 2 * n_p, N // 2 * n_s

    def f64(self):
        return self.p.astype(np.float64), None if self.s is None else self.s.astype(np.float64)

    def dev(self, a):
        return torch.as_tensor(np.ascontiguousarray(a), device="cuda")

    def ws(self, mp, op):
        n = mp.meshprop_workspace_size(op, X_CODE[self.X], self.N, self.L, self.B, int(self.fft), None)
        return torch.empty(max(n, 1), dtype=torch.uint8, device="cuda"), n


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(mp, rc):
    assert rc == 0, mp.meshprop_status_string(rc)


def _fwdprop(mp, m, u_in, u_out):
    ws, n = m.ws(mp, 0)
    if m.fft:
        rc = mp.meshprop_fwdprop_ft(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.strides), _ptr(m.p_d), m.ldp, _ptr(m.s_d),
                                    m.lds, _ptr(u_in), _ptr(u_out), m.N, 0, _ptr(ws), n, _stream())
    else:
        rc = mp.meshprop_fwdprop(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.lens), _ptr(m.shifts), _ptr(m.p_d), m.ldp,
                                 _ptr(m.s_d), m.lds, _ptr(u_in), _ptr(u_out), m.N, 0, _ptr(ws), n, _stream())
    _check(mp, rc)
    torch.cuda.synchronize()


def _fwddiff(mp, m, dp, u_in, du_in, u_out, du_out):
    ws, n = m.ws(mp, 1)
    if m.fft:
        rc = mp.meshprop_fwddiff_ft(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.strides), _ptr(m.p_d), _ptr(dp), m.ldp,
                                    _ptr(m.s_d), m.lds, _ptr(u_in), _ptr(du_in), _ptr(u_out), _ptr(du_out), m.N, 0,
                                    _ptr(ws), n, _stream())
    else:
        rc = mp.meshprop_fwddiff(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.lens), _ptr(m.shifts), _ptr(m.p_d), _ptr(dp),
                                 m.ldp, _ptr(m.s_d), m.lds, _ptr(u_in), _ptr(du_in), _ptr(u_out), _ptr(du_out), m.N,
                                 0, _ptr(ws), n, _stream())
    _check(mp, rc)
    torch.cuda.synchronize()


def _backdiff(mp, m, dp, u_out, dJdu_out, u_in, dJdu_in):
    ws, n = m.ws(mp, 2)
    if m.fft:
        rc = mp.meshprop_backdiff_ft(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.strides), _ptr(m.p_d), _ptr(dp), m.ldp,
                                     _ptr(m.s_d), m.lds, _ptr(u_out), _ptr(dJdu_out), _ptr(u_in), _ptr(dJdu_in), m.N,
                                     0, _ptr(ws), n, _stream())
    else:
        rc = mp.meshprop_backdiff(X_CODE[m.X], m.N, m.L, m.B, _ptr(m.lens), _ptr(m.shifts), _ptr(m.p_d), _ptr(dp),
                                  m.ldp, _ptr(m.s_d), m.lds, _ptr(u_out), _ptr(dJdu_out), _ptr(u_in), _ptr(dJdu_in),
                                  m.N, 0, _ptr(ws), n, _stream())
    _check(mp, rc)
    torch.cuda.synchronize()


CASES = [("mzi", "nn", 256, 20, 37), ("sym", "nn", 130, 9, 21), ("orth", "nn", 64, 12, 19),
         ("mzi", "fft", 256, 8, 23), ("sym", "fft", 64, 6, 9), ("orth", "fft", 128, 7, 17)]


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c)) for c in CASES])
def test_reference_shaped_launchers(mp, case):
    X, kind, N, L, B = case
    print(f"abi launchers: {case}", flush=True)
    m = Mesh(X, kind, N, L, B, seed=N + L + B)
    p64, s64 = m.f64()
    u64 = m.u.astype(np.complex128)
    # fwdprop
    u_in = m.dev(m.u)
    out = torch.empty_like(u_in)
    _fwdprop(mp, m, u_in, out)
    ref = O.mesh_fwd(m.g, X, p64, s64, u64)
    assert O.rel_l2(out.cpu().numpy(), ref) < TOL
    # fwddiff, with a tangent input and with du_in = NULL (zero tangent)
    dpd = m.dev(m.dp_dir.reshape(-1))
    for du in (m.du, None):
        o2, do2 = torch.empty_like(u_in), torch.empty_like(u_in)
        _fwddiff(mp, m, dpd, u_in, m.dev(du) if du is not None else None, o2, do2)
        r, dr = O.mesh_jvp(m.g, X, p64, s64, u64, m.dp_dir.astype(np.float64),
                           None if du is None else du.astype(np.complex128))
        assert O.rel_l2(o2.cpu().numpy(), r) < TOL
        assert O.rel_l2(do2.cpu().numpy(), dr) < TOL
    # backdiff in place: u and dJdu buffers are both input and output; dp accumulates
    ubuf = out.clone()
    gbuf = m.dev(m.gout)
    dp = m.dev(m.dp0)
    _backdiff(mp, m, dp, ubuf, gbuf, ubuf, gbuf)
    rdp, rgin, _ = O.mesh_vjp(m.g, X, p64, s64, out.cpu().numpy().astype(np.complex128),
                              m.gout.astype(np.complex128))
    assert O.rel_l2((dp.cpu().numpy() - m.dp0).reshape(rdp.shape), rdp) < TOL
    assert O.rel_l2(gbuf.cpu().numpy(), rgin) < TOL
    assert O.rel_l2(ubuf.cpu().numpy(), m.u) < TOL
