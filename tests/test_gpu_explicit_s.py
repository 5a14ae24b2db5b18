"""Explicit-S SR solvers (kern_dense.hip) against the oracle on host-given
sample sets, through the C ABI:

  * S = O* O^T / S - <O>* <O>^T with the CG regularisation on the diagonal
    (cg_solver.cpp:274-281), built by the MFMA SYRK;
  * "cg-single": the CG on that S (cg_solver.cpp:274-299 then 119-189),
    iterates against the oracle's matrix-free CG (the same operator);
  * "direct": S with the direct solver's multiplicative regularisation
    (direct_solver.cpp:38-55) solved exactly, against numpy on the oracle's O.

Tolerances: FP64 1e-10 of max |S| (S), 1e-9 (CG iterates, direct x);
FP32 (O in FP32, FP32 MFMA accumulation within a sample slice) 1e-4 of
max |S| and 1e-3 of max |x|.

8200 samples exercise the SYRK's many-slice layout: 32 K slices (a
multiple of 8, so each slice's tiles share one XCD), the last three of
them empty, and the fused S.p + CG step over 63+ blocks.
"""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, random_params

pytestmark = pytest.mark.gpu

KEEP = "[gpu]\nkeep_o = true\n"
SPECS = {
    "kitaev3": ini("kitaev", 3),
    "kitaev6_sym": ini("kitaev", 6, rbm="symmetry", alpha=4),
    "kitaev2_pf": ini("kitaev", 2, extra="pfaffian = true\n"),
    "toric4": ini("toric", 4, alpha=2),
}
CASES = [("kitaev3", 64), ("kitaev3", 32), ("kitaev6_sym", 64), ("kitaev6_sym", 32), ("kitaev2_pf", 64),
         ("toric4", 32)]
R1, R2, RD = 0.9, 9e-4, 1e-2


def _setup(name, prec, S=700):
    print(f"[gpu-test] explicit S {name} prec={prec}", flush=True)
    rng = np.random.default_rng(31)
    spec = SPECS[name]
    m = rbmgpu.Model(spec + KEEP)
    o = po.Oracle(spec)
    p = random_params(rng, m.P, 0.1)
    o.set_params(p)
    g = m.gpu(precision=prec)
    g.set_params(p)
    assert not g.structured()
    st = po.random_states(rng, S, m.N)
    g.load_samples(st)
    e = o.eloc(st)
    O = o.derivative(st)
    ref = o.aggregate(e, O)
    return m, o, g, rng, O, ref


def _S_ref(O, ref, n_neural, mode):
    S = O.shape[1]
    Sm = O.conj() @ O.T / S - np.outer(ref["o_mean"].conj(), ref["o_mean"])
    var = ref["o_var"]
    d = np.arange(Sm.shape[0])
    f = np.where(d < n_neural, R1, R1 + RD)
    if mode == "cg":
        Sm[d, d] += f * var + R2 * var.max()
    else:
        Sm[d, d] *= 1 + f
        Sm[d, d] += R2 * np.abs(var).max()
    return Sm


@pytest.mark.parametrize("name,prec", CASES)
def test_explicit_S_matches_oracle(name, prec):
    m, o, g, rng, O, ref = _setup(name, prec)
    g.reduce(want_o=False)
    got = g.build_S(R1, R2, RD)
    want = _S_ref(O, ref, m.n_neural, "cg")
    tol = 1e-10 if prec == 64 else 1e-4
    err = np.max(np.abs(got - want)) / np.abs(want).max()
    print(f"[gpu-test] explicit S error {err:.2e}", flush=True)
    assert err <= tol, err
    assert np.max(np.abs(got - got.conj().T)) <= tol * np.abs(want).max()


@pytest.mark.parametrize("name,prec", CASES)
def test_cg_single_matches_oracle_cg(name, prec):
    m, o, g, rng, O, ref = _setup(name, prec)
    r = g.reduce()
    F = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    x0 = 0.01 * (rng.normal(size=m.P) + 1j * rng.normal(size=m.P))
    xg, itg, _ = g.sr_solve_ex(1, F, x0, R1, R2, RD, max_iter=8, rtol=1e-300)
    xo, ito, _ = po.cg(O, ref["o_mean"], ref["o_var"], R1, R2, RD, m.n_neural, F, x0, 8, 1e-300)
    assert itg == ito == 8
    tol = 1e-9 if prec == 64 else 1e-3
    err = np.max(np.abs(xg - xo)) / np.abs(xo).max()
    print(f"[gpu-test] cg-single iterate error {err:.2e}", flush=True)
    assert err <= tol, err


@pytest.mark.parametrize("name,prec", CASES)
def test_direct_matches_numpy(name, prec):
    m, o, g, rng, O, ref = _setup(name, prec)
    g.reduce(want_o=False)
    F = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    xg, itn, rs = g.sr_solve_ex(2, F, np.zeros(m.P), R1, R2, RD)
    A = _S_ref(O, ref, m.n_neural, "direct")
    xo = np.linalg.solve(A, F)
    tol = 1e-9 if prec == 64 else 1e-3
    err = np.max(np.abs(xg - xo)) / np.abs(xo).max()
    print(f"[gpu-test] direct solve error {err:.2e}", flush=True)
    assert err <= tol, err
    assert itn == 0


@pytest.mark.parametrize("prec", [32, 64])
def test_explicit_S_many_slices_matches_oracle(prec):
    """8200 random states of the Kitaev 6x6 symmetric RBM: S with the CG
    regularisation against the oracle's O, then 6 cg-single iterations
    against the oracle's CG"""
    m, o, g, rng, O, ref = _setup("kitaev6_sym", prec, S=8200)
    g.reduce(want_o=False)
    got = g.build_S(R1, R2, RD)
    want = _S_ref(O, ref, m.n_neural, "cg")
    tol = 1e-10 if prec == 64 else 1e-4
    err = np.max(np.abs(got - want)) / np.abs(want).max()
    print(f"[gpu-test] explicit S (8200 samples) error {err:.2e}", flush=True)
    assert err <= tol, err
    F = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    x0 = np.zeros(m.P, np.complex128)
    xg, itg, _ = g.sr_solve_ex(1, F, x0, R1, R2, RD, max_iter=6, rtol=1e-300)
    xo, ito, _ = po.cg(O, ref["o_mean"], ref["o_var"], R1, R2, RD, m.n_neural, F, x0, 6, 1e-300)
    assert itg == ito == 6
    err = np.max(np.abs(xg - xo)) / np.abs(xo).max()
    print(f"[gpu-test] cg-single (8200 samples) iterate error {err:.2e}", flush=True)
    assert err <= (1e-9 if prec == 64 else 1e-3), err
