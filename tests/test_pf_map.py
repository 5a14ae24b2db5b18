"""Pins of the Pfaffian index map bs_ (pfaffian.cpp:31-85), which decides
which fs_ row and sign every entry A_ij(sigma) uses and so fixes the
meaning of every Pfaffian parameter in a .rbm file.  Both the host layer
(rbm-nqs_amd/host/setup.cpp) and the oracle build it; here both are held
to answers that do not come from either:

  * pfaffian.symmetry = 0: one identity symmetry and every site a basis
    site, so the fill writes bs(j, i) = j + N i, the transposition keeps the
    smaller magnitude i + N j for i > j, and the condensation ranks those
    magnitudes, which is the triangular index idx(i, j) + 1 of
    pfaffian.cpp:112-118 (lower triangle positive, upper negative);
  * the Kitaev 2 x 2 map at pfaffian.symmetry = 0.5, worked out by hand in
    tests/golden/pf_map_kitaev2x2_sym05.json (derivation in the file).
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _maps(spec):
    m = rbmgpu.Model(spec)
    o = po.Oracle(spec)
    return m.pf_map(), o.pf_map()


def _tri_idx(ns, i, j):
    """pfaffian::idx (pfaffian.cpp:112-118)"""
    if i < j:
        i, j = j, i
    return ns * (ns - 1) // 2 - (ns - j) * (ns - j - 1) // 2 + i - j - 1


@pytest.mark.parametrize("model,n", [("kitaev", 2), ("kitaev", 3), ("toric", 3)])
def test_pf_map_symmetry0_closed_form(model, n):
    spec = ini(model, n, extra="pfaffian = true\npfaffian.symmetry = 0\n")
    (bm, rm), (bo, ro) = _maps(spec)
    N = bm.shape[0]
    want = np.zeros((N, N), dtype=np.int64)
    for i in range(N):
        for j in range(i):
            want[i, j] = _tri_idx(N, i, j) + 1
            want[j, i] = -want[i, j]
    assert rm == ro == N * (N - 1) // 2
    assert np.array_equal(bm, want)
    assert np.array_equal(bo, want)


def test_pf_map_kitaev2x2_symmetry_half_fixture():
    d = json.load(open(os.path.join(GOLDEN, "pf_map_kitaev2x2_sym05.json")))
    want = np.array(d["bs"], dtype=np.int64)
    (bm, rm), (bo, ro) = _maps(d["ini"])
    assert rm == ro == d["rows"]
    assert np.array_equal(bm, want)
    assert np.array_equal(bo, want)


@pytest.mark.parametrize("model,n,sym", [("kitaev", 3, 0.5), ("kitaev", 4, 0.5), ("toric", 3, 1)])
def test_pf_map_host_equals_oracle(model, n, sym):
    """larger lattices: the two restatements agree, the map is skew with a
    zero diagonal, and every row 1..rows of fs_ is used"""
    (bm, rm), (bo, ro) = _maps(ini(model, n, extra=f"pfaffian = true\npfaffian.symmetry = {sym}\n"))
    assert rm == ro and np.array_equal(bm, bo)
    assert np.array_equal(bm, -bm.T) and not np.any(np.diag(bm))
    assert set(np.abs(bm[np.tril_indices(bm.shape[0], -1)]).tolist()) == set(range(1, rm + 1))
