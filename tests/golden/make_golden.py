"""Writes the golden fixtures under tests/golden/ by running the reference CPU engine.

The reference engine ``c/dot.cpp`` is compiled from its own source by ``oracle/Makefile``
into ``oracle/_ref/libdot_ref.so`` (run ``make -C oracle`` first, in a container where
/root/reference exists).  This script feeds it seeded inputs and stores inputs and
outputs as small ``.npz`` files (data only).  The engine takes T directly, so the
fixtures pin the propagation engine (modes 0/1/2, packed NN geometry); the crossing
formulas are pinned separately by closed forms and finite differences.

Mode 2 (VJP) is run with threads=1: c/dot.cpp:135 and :279-285 update shared
accumulators inside the OpenMP region.

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import meshprop_oracle as O  # noqa: E402
from oracle import ref_engine  # noqa: E402


def _case(name, geom, X, p, s, u, dp, g_target_seed, threads=8):
    T = O.slot_T(geom, X, p, s)
    dT = O.slot_dT_dir(geom, X, p, s, dp)
    out0 = ref_engine.run(geom, T, u, 0, threads=threads)
    out1, dout1 = ref_engine.run(geom, T, u, 1, dT_slot=dT, threads=threads)
    rng = np.random.RandomState(g_target_seed)
    y = O.unit_cn(rng, u.shape[0], geom.N)
    g = 2 * (out0 - y)
    rec2, gin2, dJdT2 = ref_engine.run(geom, T, out0, 2, dv=g, threads=1)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(
        path, N=geom.N, L=geom.L, lens=geom.lens, shifts=geom.shifts, X=X,
        p=p, s=(np.zeros(0) if s is None else s), dp=dp, T=T, dT=dT, u_in=u, y=y,
        fwd_out=out0, jvp_out=out1, jvp_dout=dout1, g_out=g,
        vjp_u_in=rec2, vjp_g_in=gin2, vjp_dJdT=dJdT2)
    print("wrote", path)


def _packed_case(name, N, lens, shifts, perms, X, p_packed, u, dp_packed, ph, dph, pos, seed):
    """A packed-layout mesh (optionally with permutations and a phase screen) through the
    reference engine, modes 0/1/2 (mode 2 with the inverse permutations, mesh.py:381)."""
    T = O.crossing_T(X, p_packed).transpose(1, 2, 0)
    dTk = O.crossing_dT(X, p_packed)  # (n, n_p, 2, 2)
    dT = np.einsum("nkij,nk->ijn", dTk, dp_packed)
    out0 = ref_engine.run_packed(N, lens, shifts, T, u, 0, ph=ph, perms=perms, pos=pos, threads=4)
    out1, dout1 = ref_engine.run_packed(N, lens, shifts, T, u, 1, dT=dT, dv=None, ph=ph, dph=dph, perms=perms,
                                        pos=pos, threads=4)
    rng = np.random.RandomState(seed)
    y = O.unit_cn(rng, u.shape[0], N)
    g = 2 * (out0 - y)
    perms_bk = None if perms is None else np.array([np.argsort(q) for q in perms])
    rec2, gin2, dJdT2, dph2 = ref_engine.run_packed(N, lens, shifts, T, out0, 2, dv=g, ph=ph, perms=perms_bk,
                                                    pos=pos, threads=1)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(
        path, N=N, lens=np.asarray(lens), shifts=np.asarray(shifts),
        perms=(np.zeros((0, 0), int) if perms is None else np.asarray(perms)), X=X, p=p_packed, dp=dp_packed,
        ph=ph, dph=dph, pos=pos, u_in=u, y=y, fwd_out=out0, jvp_out=out1, jvp_dout=dout1, g_out=g,
        vjp_u_in=rec2, vjp_g_in=gin2, vjp_dJdT=dJdT2, vjp_dph=dph2)
    print("wrote", path)


def main():
    if not ref_engine.available():
        raise SystemExit("build the reference engine first: make -C oracle")

    # Config C1: Clements MZI, N=16, L=16, phases U[0, 2pi), 64 unit-norm CN(0,1) vectors, seed 1.
    rng = np.random.RandomState(1)
    geom = O.clements(16)
    p = rng.uniform(0, 2 * np.pi, size=(16, 8, 2))
    u = O.unit_cn(rng, 64, 16)
    dp = rng.standard_normal((16, 8, 2))
    dp[~O.active_mask(geom)] = 0
    dp /= np.linalg.norm(dp)
    _case("c1_clements16", geom, "mzi", p, None, u, dp, 101)

    # Random NN geometries (gpu/test.py:41 style), splitter errors, N not a power of two.
    for k, (N, L, B) in enumerate([(10, 7, 5), (24, 13, 17), (38, 20, 33), (64, 9, 40)]):
        rng = np.random.RandomState(200 + k)
        geom = O.random_geometry_nn(rng, N, L)
        p = rng.standard_normal((L, N // 2, 2))
        s = 0.1 * rng.standard_normal((L, N // 2, 2))
        u = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
        dp = rng.standard_normal((L, N // 2, 2))
        _case(f"nn_random_{N}", geom, "mzi", p, s, u, dp, 300 + k)

    # Symmetric crossing, Clements N=12 (the engine takes any T).
    rng = np.random.RandomState(400)
    geom = O.clements(12)
    p = rng.uniform(0, 2 * np.pi, size=(12, 6, 2))
    s = 0.05 * rng.standard_normal((12, 6, 1))
    u = O.unit_cn(rng, 9, 12)
    dp = rng.standard_normal((12, 6, 2))
    _case("sym_clements12", geom, "sym", p, s, u, dp, 401)


    # Packed layout with permutations and the phase screen: a ButterflyNetwork geometry
    # (butterfly.py:46-51, N-1 layers with stride gathers) with the output screen, and a
    # Clements mesh with the input screen.
    rng = np.random.RandomState(500)
    N = 16
    lens, shifts, perms = O.butterfly_perms(N)
    n_cr = sum(lens)
    p = rng.uniform(0, 2 * np.pi, size=(n_cr, 2))
    u = O.unit_cn(rng, 11, N)
    dp = rng.standard_normal((n_cr, 2))
    ph = rng.uniform(0, 2 * np.pi, N)
    dph = rng.standard_normal(N)
    _packed_case("butterfly16_screen_out", N, lens, shifts, perms, "mzi", p, u, dp, ph, dph, 1, 501)
    rng = np.random.RandomState(600)
    N = 12
    geom = O.clements(N)
    lens, shifts = list(geom.lens), list(geom.shifts)
    n_cr = sum(lens)
    p = rng.uniform(0, 2 * np.pi, size=(n_cr, 2))
    u = O.unit_cn(rng, 7, N)
    dp = rng.standard_normal((n_cr, 2))
    ph = rng.uniform(0, 2 * np.pi, N)
    dph = rng.standard_normal(N)
    _packed_case("clements12_screen_in", N, lens, shifts, None, "mzi", p, u, dp, ph, dph, 0, 601)


if __name__ == "__main__":
    main()
