"""CPU tests of the host layer: the C ABI libraries load and export every
declared symbol, and the host-built lattice / Hamiltonian / machine tables
equal the oracle's (which are pinned to the reference's golden data)."""
import os
import re

import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SPECS = [ini("kitaev", 2), ini("kitaev", 3), ini("kitaev", 3, rbm="symmetry", alpha=2),
         ini("kitaev", 4, alpha=2), ini("kitaev", 6, rbm="symmetry", alpha=4),
         ini("toric", 2), ini("toric", 4, alpha=2), ini("toric", 3, rbm="symmetry", alpha=2)]


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\w[\w\s\*]*?\b(rbm\w+)\s*\(", text, flags=re.M))


def test_libraries_export_declared_symbols():
    g, h = rbmgpu.libs()
    declared = _declared("rbmgpu.h") | _declared("rbmhost.h")
    assert len(declared) >= 25
    for name in declared:
        lib = g if name.startswith("rbmgpu_") else h
        assert hasattr(lib, name), name
    listed = set(rbmgpu.exported_symbols()["gpu"]) | set(rbmgpu.exported_symbols()["host"])
    assert listed == declared


@pytest.mark.parametrize("spec", SPECS)
def test_host_tables_equal_oracle(spec):
    m = rbmgpu.Model(spec)
    o = po.Oracle(spec)
    assert (m.N, m.P, m.n_vb, m.n_alpha, m.n_symm) == (o.N, o.P, o.n_vb, o.n_alpha, o.n_symm)
    assert np.array_equal(m.bonds(), o.bonds())
    assert np.array_equal(m.perms(), o.symmetry())
    terms_o = o.terms()
    terms_m = m.terms()
    assert len(terms_o) == len(terms_m)
    for (so, mo), (sm, mm, g) in zip(terms_o, terms_m):
        assert np.array_equal(so, sm)
        assert np.array_equal(mo, mm)
        assert g == 0
    assert abs(m.exact_energy() - o.exact_energy()) < 1e-14


def test_pfaffian_layout_matches_oracle():
    spec = ini("kitaev", 4, alpha=2, extra="pfaffian = true\n")
    m = rbmgpu.Model(spec)
    o = po.Oracle(spec)
    assert m.pf_rows * 4 == o.n_pf <= 124  # <= 31 rows x 4 at symmetry 0.5 (N = 32)
    assert m.P == o.P == 2144 + o.n_pf


def test_evaluation_observables():
    m = rbmgpu.Model(ini("kitaev", 3, extra="[model]\nhelper_strength = 0.5\n"))
    assert m.n_terms == 27 + 9
    m.use_evaluation(hexagons=True)
    assert m.n_terms == 27 + 9 and m.n_groups == 10
    groups = [g for _, _, g in m.terms()]
    assert groups == [0] * 27 + list(range(1, 10))


def test_initial_weights_reproducible():
    m = rbmgpu.Model(ini("toric", 2))
    a = m.initial_weights(421390484)
    b = m.initial_weights(421390484)
    assert np.array_equal(a, b)
    assert abs(a.real.std() - 1e-4) < 3e-5 and abs(a.imag.std() - 1e-4) < 3e-5


def test_ini_errors_are_reported():
    with pytest.raises(RuntimeError, match="Model Type"):
        rbmgpu.Model("[model]\ntype = ising\n")
    with pytest.raises(RuntimeError, match="unrecognised option"):
        rbmgpu.Model("[model]\nnot_a_key = 3\n")


def test_eval_samples_rescale_warmup():
    # metropolis_sampler.cpp:289-293
    m = rbmgpu.Model(ini("toric", 2, extra="[sampler]\nn_samples = 1000\nmetropolis.n_warmup_steps = 100\n"))
    m.set_n_samples(5000)
    assert m.n_samples == 5000 and m.n_warmup == 500


@pytest.mark.parametrize("model,n", [("kitaev", 2), ("kitaev", 3), ("toric", 3), ("toric", 4)])
def test_cluster_catalogue_matches_oracle(model, n):
    """learned-proposal clusters: the first lattice loop through each site"""
    spec = ini(model, n)
    assert np.array_equal(rbmgpu.Model(spec).clusters(), po.Oracle(spec).clusters())


def test_massive_step_guard():
    """abstract_optimizer.cpp:157: an update with max |dw| > 30 is refused
    (the driver then returns rc 55); non-finite entries are refused too."""
    dw = np.zeros(100, np.complex128)
    dw[7] = 29.9 + 0.1j
    assert rbmgpu.step_ok(dw)[0]
    assert abs(rbmgpu.step_ok(dw)[1] - abs(dw[7])) < 1e-15
    dw[7] = 30.0
    assert rbmgpu.step_ok(dw)[0]
    dw[7] = 30.0 + 1e-6
    assert not rbmgpu.step_ok(dw)[0]
    dw[7] = 0
    assert rbmgpu.step_ok(dw) == (True, 0.0)
    for bad in (np.nan, np.inf, complex(0, np.nan), complex(-np.inf, 0)):
        dw[3] = bad
        assert not rbmgpu.step_ok(dw)[0]
    dw[3] = 0
    dw[99] = np.nan  # the last entry counts as well
    assert not rbmgpu.step_ok(dw)[0]
