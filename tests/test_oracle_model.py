"""Pins the oracle's Hamiltonian and local energy to exact diagonalisation
(notebooks/KitaevED.ipynb cells 10 and 13), the closed form of
kitaev.cpp:66-100 and the toric-code ground energy."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh

import pyoracle as po
from helpers import ini, random_params


def full_hamiltonian(o):
    """Sparse 2^N x 2^N H from the oracle's local terms; basis index bit i is
    site i with bit 1 <=> sigma_i = -1 (the ED notebook's ordering)."""
    N = o.N
    dim = 1 << N
    idx = np.arange(dim, dtype=np.int64)
    rows, cols, vals = [], [], []
    for sites, mat in o.terms():
        loc = np.zeros(dim, dtype=np.int64)
        for k, s in enumerate(sites):
            loc |= ((idx >> s) & 1) << k
        d = mat.shape[0]
        for j in range(d):
            v = mat[loc, j]
            nz = v != 0
            if not nz.any():
                continue
            x = loc[nz] ^ j
            tgt = idx[nz].copy()
            for k, s in enumerate(sites):
                tgt ^= ((x >> k) & 1) << s
            rows.append(idx[nz])
            cols.append(tgt)
            vals.append(v[nz])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(dim, dim))


@pytest.mark.parametrize("n,expected", [(2, -0.8660254037844461), (3, -0.7939723678960895)])
def test_kitaev_ed_energy(n, expected):
    o = po.Oracle(ini("kitaev", n))
    H = full_hamiltonian(o)
    assert abs(H - H.getH()).max() < 1e-14
    w = eigsh(H, k=1, which="SA")[0][0]
    assert abs(w / o.N - expected) < 1e-10


def test_kitaev_closed_form_matches_ed():
    assert abs(po.Oracle(ini("kitaev", 3)).exact_energy() - (-0.7939723678960895)) < 1e-12
    # values evaluated from kitaev.cpp:66-100 for the larger lattices
    assert abs(po.Oracle(ini("kitaev", 4)).exact_energy() - (-0.79426274578121)) < 1e-12
    assert abs(po.Oracle(ini("kitaev", 6)).exact_energy() - (-0.78823570803603)) < 1e-12


def test_toric_ground_energy():
    o = po.Oracle(ini("toric", 2))
    H = full_hamiltonian(o)
    w = eigsh(H, k=1, which="SA")[0][0]
    assert abs(w / o.N - (-1.0)) < 1e-10


@pytest.mark.parametrize("spec", [ini("kitaev", 2), ini("kitaev", 2, rbm="symmetry", alpha=2),
                                  ini("toric", 2), ini("toric", 2, rbm="symmetry", alpha=2)])
def test_eloc_equals_H_psi_over_psi(spec):
    """E_loc(s) = sum_s' H_ss' psi(s')/psi(s) from the dense wavefunction."""
    rng = np.random.default_rng(3)
    o = po.Oracle(spec)
    o.set_params(random_params(rng, o.P, 0.3))
    N = o.N
    states = np.arange(1 << N, dtype=np.uint64).reshape(-1, 1)
    lp = o.logpsi(states)
    psi = np.exp(lp - lp.real.max())
    H = full_hamiltonian(o)
    expect = (H @ psi) / psi
    got = o.eloc(states)
    assert np.allclose(got, expect, rtol=1e-11, atol=1e-11)


def test_exact_expectation_is_rayleigh_quotient():
    rng = np.random.default_rng(4)
    o = po.Oracle(ini("kitaev", 2))
    o.set_params(random_params(rng, o.P, 0.2))
    states = np.arange(1 << o.N, dtype=np.uint64).reshape(-1, 1)
    lp = o.logpsi(states)
    psi = np.exp(lp - lp.real.max())
    H = full_hamiltonian(o)
    rq = np.vdot(psi, H @ psi) / np.vdot(psi, psi)
    assert abs(o.exact_expectation() - rq) < 1e-12
