"""Host-side bridge logic (no GPU): StructuredMeshNetwork-like objects onto the GPU geometry.

The permuted butterfly of butterfly.py:46-51 maps onto FFT strides with a slot order such that
the FFT-geometry product (the oracle's Geometry / fwd) equals the packed product with the
permutations (structured_dense, pinned to c/dot.cpp by the golden fixtures); grid_phase mirrors
mesh.py:513-522.
"""
import numpy as np
import pytest

from oracle import meshprop_oracle as O


def test_butterfly_strides_and_slots_reproduce_the_permuted_mesh():
    from meshes_amd.bridge import strides_from_perms
    rng = np.random.RandomState(3)
    for N in (4, 8, 16, 32):
        lens, shifts, perms = O.butterfly_perms(N)
        strides, slots = strides_from_perms(N, lens, shifts, perms)
        assert list(strides[:7]) == [0, 1, 0, 2, 0, 1, 0][:N - 1]
        p = rng.uniform(0, 2 * np.pi, (sum(lens), 2))
        T = O.crossing_T("mzi", p).transpose(1, 2, 0)
        Tslot = np.zeros((N - 1, N // 2, 2, 2), complex)
        ind = 0
        for n in range(N - 1):
            Tslot[n, slots[n]] = T[:, :, ind:ind + lens[n]].transpose(2, 0, 1)
            ind += lens[n]
        g = O.Geometry(N, N - 1, strides=strides)
        U_fft = O.dense_matrix(g, Tslot)
        U_ref = O.structured_dense(N, lens, shifts, T, perms=perms)
        assert O.rel_l2(U_fft, U_ref) < 1e-13


def test_non_butterfly_permutations_are_rejected():
    from meshes_amd.bridge import strides_from_perms
    N = 8
    lens, shifts = [N // 2] * 2, [0] * 2
    bad = [np.array([0, 3, 1, 2, 4, 5, 6, 7]), None, None]
    with pytest.raises(ValueError):
        strides_from_perms(N, lens, shifts, bad)
    unrestored = [np.array([0, 2, 1, 3, 4, 6, 5, 7]), None, None]
    with pytest.raises(ValueError):
        strides_from_perms(N, lens, shifts, unrestored)


class _X:
    def __init__(self, n_phase=2, n_splitter=2):
        self.n_phase, self.n_splitter = n_phase, n_splitter


def test_grid_phase_matches_the_slot_layout():
    from meshes_amd.bridge import grid_phase
    N = 6
    g = O.clements(N)
    lens, shifts = list(g.lens), list(g.shifts)
    n_cr = sum(lens)

    class Net:
        pass
    net = Net()
    net.N, net.L, net.lens, net.shifts, net.is_phase = N, len(lens), lens, shifts, True
    net.X = _X()
    net.p_phase = np.arange(n_cr * 2 + N, dtype=float)
    net.p_splitter = 0.0
    grid = grid_phase(net)
    assert grid.shape == (N, 3, 2)
    mask = O.active_mask(g)
    assert np.all(np.isnan(grid[~mask]))
    assert np.array_equal(grid[mask].reshape(-1), np.arange(n_cr * 2, dtype=float))


def test_permuted_mesh_segments():
    """A permuted mesh splits at every layer with a gather before it (the gathers before layer 0
    and after the last run outside the segments)."""
    from meshes_amd.permuted import segment_bounds
    P = np.arange(4)
    assert segment_bounds(5, [None] * 6) == [0, 5]
    assert segment_bounds(5, [P, None, P, None, None, P]) == [0, 2, 5]
    assert segment_bounds(3, [None, P, P, None]) == [0, 1, 2, 3]
