"""ctypes driver for the compiled reference CPU engine (``oracle/_ref/libdot_ref.so``).

TEST INFRASTRUCTURE ONLY.  The library is built by ``oracle/Makefile`` from the
reference's own source ``c/dot.cpp``; this module converts between the slot layout used
by the GPU path and the packed layout the reference engine takes, following
``mesh.py:97-113`` (``meshdot_helper_c``) and ``mesh.py:343-392`` (``_dot_pars``,
``dot``, ``dot_vjp``).  It never imports the reference Python package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libdot_ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB_PATH)
        for name in ("meshdot_helper32", "meshdot_helper64"):
            fn = getattr(lib, name)
            # c/dot.cpp:310-325
            fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 10 + [ctypes.c_int] * 4
            fn.restype = None
        _lib = lib
    return _lib


def pack_slots(geom, T_slot):
    """Slot layout [L, N/2, 2, 2] -> packed (2, 2, n_cr) with inds = cumsum(lens) (mesh.py:270-272)."""
    cols = []
    for l in range(geom.L):
        m = geom.pairs(l)[0]
        cols.append(T_slot[l, m])
    packed = np.concatenate(cols, axis=0) if cols else np.zeros((0, 2, 2), np.complex128)
    return np.ascontiguousarray(packed.transpose(1, 2, 0))


def unpack_slots(geom, packed):
    out = np.zeros((geom.L, geom.N // 2, 2, 2), dtype=packed.dtype)
    k = 0
    for l in range(geom.L):
        m = geom.pairs(l)[0]
        out[l, m] = packed[:, :, k:k + len(m)].transpose(2, 0, 1)
        k += len(m)
    return out


def run(geom, T_slot, v, mode, dT_slot=None, dv=None, threads=1, dtype=np.complex128):
    """Run meshdot_helper on fields v given as (B, N) rows.

    mode 0: returns v_out.  mode 1: returns (v_out, dv_out).  mode 2 (dv = dJ/dv_out):
    returns (v_in reconstructed, dJdv_in, dJdT in slot layout).  No phase screen
    (pos = -1) and no permutations, as in the GPU path (gpu/mesh.py:65)."""
    assert geom.kind == "nn"
    lib = _load()
    is64 = dtype == np.complex128
    fn = lib.meshdot_helper64 if is64 else lib.meshdot_helper32
    ftype = np.float64 if is64 else np.float32
    N, L = geom.N, geom.L
    B = v.shape[0]
    T = np.ascontiguousarray(pack_slots(geom, T_slot), dtype=dtype)
    nT = T.shape[2]
    if mode == 2:
        dT = np.zeros_like(T)
    elif mode == 1:
        dT = np.ascontiguousarray(pack_slots(geom, dT_slot), dtype=dtype)
    else:
        dT = np.zeros((0, 0, 0), dtype=dtype)
    vv = np.ascontiguousarray(np.asarray(v).T, dtype=dtype)  # (N, B) C order, batch contiguous
    if mode == 0:
        dvv = np.zeros((0, 0), dtype=dtype)
    else:
        dvv = np.zeros_like(vv) if dv is None else np.ascontiguousarray(np.asarray(dv).T, dtype=dtype)
    ph = np.zeros(N, dtype=ftype)
    dph = np.zeros(N, dtype=ftype)
    inds = np.ascontiguousarray(np.cumsum([0] + list(geom.lens)), dtype=np.int32)
    lens = np.ascontiguousarray(geom.lens, dtype=np.int32)
    shifts = np.ascontiguousarray(geom.shifts, dtype=np.int32)
    perms = np.zeros((0, 0), dtype=np.int32)
    fn(N, B, nT, L, T.ctypes.data, dT.ctypes.data, ph.ctypes.data, dph.ctypes.data,
       vv.ctypes.data, dvv.ctypes.data, inds.ctypes.data, lens.ctypes.data, shifts.ctypes.data,
       perms.ctypes.data, 0, -1, int(mode), int(threads))
    if mode == 0:
        return vv.T.copy()
    if mode == 1:
        return vv.T.copy(), dvv.T.copy()
    return vv.T.copy(), dvv.T.copy(), unpack_slots(geom, dT)


def run_packed(N, lens, shifts, T, v, mode, dT=None, dv=None, ph=None, dph=None, perms=None, pos=-1, threads=1,
               dtype=np.complex128):
    """meshdot_helper on the packed layout with permutations and the phase screen
    (mesh.py:343-392 argument preparation).  T, dT: (2, 2, nT); v, dv: (B, N) rows; perms:
    (L+1, N) gathers (for mode 2, the inverse permutations, as mesh.py:381 passes them).
    Returns what the engine leaves in (v, dv, dT, dph) for the mode, like structured_run."""
    lib = _load()
    is64 = dtype == np.complex128
    fn = lib.meshdot_helper64 if is64 else lib.meshdot_helper32
    ftype = np.float64 if is64 else np.float32
    L = len(lens)
    B = v.shape[0]
    TT = np.ascontiguousarray(T, dtype=dtype)
    nT = TT.shape[2]
    if mode == 2:
        dTT = np.zeros_like(TT)
    elif mode == 1:
        dTT = np.ascontiguousarray(dT, dtype=dtype)
    else:
        dTT = np.zeros((0, 0, 0), dtype=dtype)
    vv = np.ascontiguousarray(np.asarray(v).T, dtype=dtype)
    if mode == 0:
        dvv = np.zeros((0, 0), dtype=dtype)
    else:
        dvv = np.zeros_like(vv) if dv is None else np.ascontiguousarray(np.asarray(dv).T, dtype=dtype)
    phh = np.ascontiguousarray(np.zeros(N) if ph is None else ph, dtype=ftype)
    dphh = np.ascontiguousarray(np.zeros(N) if dph is None else dph, dtype=ftype)
    inds = np.ascontiguousarray(np.cumsum([0] + list(lens)), dtype=np.int32)
    ln = np.ascontiguousarray(lens, dtype=np.int32)
    sh = np.ascontiguousarray(shifts, dtype=np.int32)
    pp = np.ascontiguousarray(np.zeros((0, 0)) if perms is None else perms, dtype=np.int32)
    fn(N, B, nT, L, TT.ctypes.data, dTT.ctypes.data, phh.ctypes.data, dphh.ctypes.data, vv.ctypes.data,
       dvv.ctypes.data, inds.ctypes.data, ln.ctypes.data, sh.ctypes.data, pp.ctypes.data, int(pp.size), int(pos),
       int(mode), int(threads))
    if mode == 0:
        return vv.T.copy()
    if mode == 1:
        return vv.T.copy(), dvv.T.copy()
    return vv.T.copy(), dvv.T.copy(), dTT, dphh
