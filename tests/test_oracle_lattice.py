"""Pins the oracle's lattice and Hamiltonian tables to the reference's own
known answers: test/test_lattice.cpp and notebooks/KitaevED.ipynb."""
import numpy as np
import pytest

import pyoracle as po
from helpers import ini

# notebooks/KitaevED.ipynb cell 4 (2x2) and cell 6 (3x3): bonds and hexagons
BONDS2 = [([0, 1], 0), ([0, 3], 1), ([0, 5], 2), ([2, 3], 0), ([2, 1], 1), ([2, 7], 2),
          ([4, 5], 0), ([4, 7], 1), ([4, 1], 2), ([6, 7], 0), ([6, 5], 1), ([6, 3], 2)]
PLAQS2 = [[5, 0, 1, 2, 7, 6], [7, 2, 3, 0, 5, 4], [1, 4, 5, 6, 3, 2], [3, 6, 7, 4, 1, 0]]
BONDS3 = [([0, 1], 0), ([0, 5], 1), ([0, 13], 2), ([2, 3], 0), ([2, 1], 1), ([2, 15], 2),
          ([4, 5], 0), ([4, 3], 1), ([4, 17], 2), ([6, 7], 0), ([6, 11], 1), ([6, 1], 2),
          ([8, 9], 0), ([8, 7], 1), ([8, 3], 2), ([10, 11], 0), ([10, 9], 1), ([10, 5], 2),
          ([12, 13], 0), ([12, 17], 1), ([12, 7], 2), ([14, 15], 0), ([14, 13], 1), ([14, 9], 2),
          ([16, 17], 0), ([16, 15], 1), ([16, 11], 2)]
PLAQS3 = [[13, 0, 1, 2, 15, 14], [15, 2, 3, 4, 17, 16], [17, 4, 5, 0, 13, 12],
          [1, 6, 7, 8, 3, 2], [3, 8, 9, 10, 5, 4], [5, 10, 11, 6, 1, 0],
          [7, 12, 13, 14, 9, 8], [9, 14, 15, 16, 11, 10], [11, 16, 17, 12, 7, 6]]


@pytest.mark.parametrize("n,bonds,plaqs", [(2, BONDS2, PLAQS2), (3, BONDS3, PLAQS3)])
def test_honeycomb_bonds_and_hexagons_match_ed_notebook(n, bonds, plaqs):
    o = po.Oracle(ini("kitaev", n))
    got = [([int(a), int(b)], int(t)) for a, b, t in o.bonds()]
    assert got == bonds
    assert o.hexagons().tolist() == plaqs


def test_honeycomb_neighbours_kat():
    # test_lattice.cpp:242-255
    o = po.Oracle(ini("kitaev", 3))
    assert o.nns(8, 3).tolist() == [9, 7, 3]
    assert o.nns(9, 3).tolist() == [8, 10, 14]
    assert o.nns(0, 3).tolist() == [1, 5, 13]


def _mat(rows, cols):
    # Eigen column-major mat(i) = i
    return np.arange(rows * cols).reshape(cols, rows).T


@pytest.mark.parametrize("N,Nb,sh", [(5, 0, 0), (5, 0, 2), (5, 3, 0), (5, 3, 2)])
def test_bravais_up_down_tables(N, Nb, sh):
    # test_lattice.cpp:53-223
    R = Nb if Nb else N
    mat = _mat(R, N)
    for s in range(1, R):
        for i in range(R):
            for j in range(N):
                up, down = po.bravais_updown(N, Nb, sh, mat[i, j], 0, s)
                assert up == mat[(i + s) % R, j]
                assert down == mat[(i - s + R) % R, j]
    for s in range(1, N):
        for i in range(R):
            for j in range(N):
                up, down = po.bravais_updown(N, Nb, sh, mat[i, j], 1, s)
                inew_up = (R + i - (sh if j + s >= N else 0)) % R
                inew_dn = (i + (sh if j - s < 0 else 0)) % R
                assert up == mat[inew_up, (j + s) % N]
                assert down == mat[inew_dn, (j - s + N) % N]


def test_honeycomb_symmetry_group():
    # 0.5 symmetry: all translations x {id, rot180} (honeycomb.cpp:84-134)
    o = po.Oracle(ini("kitaev", 3, rbm="symmetry"))
    perms = o.symmetry()
    assert perms.shape == (18, 18)
    for p in perms:
        assert sorted(p.tolist()) == list(range(18))
    assert perms[0].tolist() == list(range(18))
    assert perms[1].tolist() == list(range(17, -1, -1))
    # the permutations map bonds to bonds of the same type
    bonds = {(int(a), int(b)): int(t) for a, b, t in o.bonds()}
    for p in perms:
        for (a, b), t in bonds.items():
            pa, pb = p[a], p[b]
            assert bonds.get((pa, pb), bonds.get((pb, pa))) == t


def test_toric_plaquettes():
    o = po.Oracle(ini("toric", 3))
    pl = o.plaqs()
    assert pl.shape == (18, 5)
    assert pl[:, 4].tolist() == [0, 1] * 9
    # every edge belongs to exactly 2 plaquettes and 2 vertices (torus)
    for typ in (0, 1):
        cnt = np.bincount(pl[pl[:, 4] == typ, :4].ravel(), minlength=18)
        assert (cnt == 2).all()
