"""GPU parity of the Pfaffian factor (pfaffian.cpp) against the FP64 oracle.

One Markov chain per wavefront; the inverse of A(sigma) lives in registers
and is updated by Sherman-Morrison-Woodbury, with a fresh Gauss-Jordan
inverse at chain start and every sampler.pfaffian_refresh accepted moves.

Tolerances: FP64 1e-10 relative for log psi, E_loc and O (the inverse of
A(sigma) comes from a different elimination order than the oracle's, so
agreement is to rounding, amplified by the condition number of A); FP32 as
stated per test.  Imaginary parts of log psi are compared modulo 2 pi.
"""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, random_params

pytestmark = pytest.mark.gpu

PF = "pfaffian = true\n"
KIT2 = ini("kitaev", 2, extra=PF + "[sampler]\nn_samples = 256\nmetropolis.n_chains = 8\npfaffian_refresh = 4\n")
KIT3 = ini("kitaev", 3, extra=PF + "[sampler]\nn_samples = 64\nmetropolis.n_chains = 4\npfaffian_refresh = 8\n")
KIT4 = ini("kitaev", 4, alpha=2,
           extra=PF + "[sampler]\nn_samples = 128\nmetropolis.n_chains = 64\npfaffian_refresh = 16\n")
TOR3 = ini("toric", 3, extra=PF + "[sampler]\nn_samples = 64\nmetropolis.n_chains = 8\npfaffian_refresh = 8\n")
SPECS = {"kit2": KIT2, "kit3": KIT3, "kit4": KIT4, "tor3": TOR3}


def _wrap(x):
    return (x + np.pi) % (2 * np.pi) - np.pi


def _setup(spec, prec, std=0.3, seed=3):
    print(f"[gpu-test pf] spec={spec.split()[3:6]} prec={prec} std={std}", flush=True)
    rng = np.random.default_rng(seed)
    m = rbmgpu.Model(spec)
    assert m.P > m.n_neural, "spec has no Pfaffian block"
    o = po.Oracle(spec)
    p = random_params(rng, m.P, std)
    o.set_params(p)
    g = m.gpu(precision=prec)
    g.set_params(p)
    return m, o, g, rng


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_pf_logpsi(name, prec):
    m, o, g, rng = _setup(SPECS[name], prec)
    states = po.random_states(rng, 32, m.N)
    a = g.eval_logpsi(states)
    b = o.logpsi(states)
    # FP32: Parlett-Reid over N = 32 pivots in FP32 -> stated 1e-4
    tol = 1e-10 if prec == 64 else 1e-4
    scale = 1 + np.abs(b.real)
    assert np.all(np.abs(a.real - b.real) <= tol * scale), np.max(np.abs(a.real - b.real) / scale)
    assert np.all(np.abs(_wrap(a.imag - b.imag)) <= tol * scale)


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_pf_psi_ratio(name, prec):
    """1, 2, 4 and 6 flips: the Woodbury ratio of pfaffian.cpp:299-397 with
    Pf(Cinv') in closed form (m <= 2) or by Parlett-Reid (m > 2)."""
    m, o, g, rng = _setup(SPECS[name], prec)
    for nf in (1, 2, 4, 6, 1, 2):
        s = po.random_states(rng, 1, m.N)
        f = rng.choice(m.N, nf, replace=False)
        a = g.psi_ratio(s, f)
        b = o.psi_ratio(s, f)
        # FP32: inverse of A in FP32 (condition number ~1e2) -> stated 2e-3
        tol = 1e-9 if prec == 64 else 2e-3
        assert abs(a - b) <= tol * abs(b), (nf, a, b)


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_pf_eloc_and_O(name, prec):
    m, o, g, rng = _setup(SPECS[name], prec, std=0.3 if prec == 64 else 0.2)
    states = po.random_states(rng, 48, m.N)
    g.load_samples(states)
    st, obs, O, _ = g.get_samples(states=True, obs=True, O=True)
    e_ref = o.eloc(states)
    O_ref = o.derivative(states)
    tol = 1e-9 if prec == 64 else 2e-3
    escale = np.abs(e_ref).max() + 1
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= tol * escale
    # the Pfaffian block separately (d ln Pf / d fs, pfaffian.cpp:399-415)
    n = m.n_neural
    pscale = np.abs(O_ref[n:]).max()
    assert np.max(np.abs(O[n:] - O_ref[n:])) <= tol * pscale
    assert np.max(np.abs(O[:n] - O_ref[:n])) <= tol * np.abs(O_ref[:n]).max()


@pytest.mark.parametrize("name", ["kit2", "kit4", "tor3"])
def test_pf_sampler_trajectory_fp64(name):
    """Same Philox streams, Woodbury updates and refreshes: the FP64 chains
    follow the oracle's state for state."""
    m, o, g, rng = _setup(SPECS[name], 64, std=0.1)
    g.sample_epoch(epoch=2)
    st, obs, _, acc = g.get_samples(states=True, obs=True, O=False, chain_acc=True)
    S = m.n_samples
    st_o, e_o, _, acc_o = o.sample(seed=421390484, epoch=2, n_samples=S, n_chains=m.n_chains)
    same = np.all(st == st_o, axis=1)
    assert same.mean() == 1.0, f"{(~same).sum()} of {S} samples differ"
    assert np.max(np.abs(obs[:, 0] - e_o)) <= 1e-9 * (np.abs(e_o).max() + 1)
    assert np.allclose(acc, acc_o, rtol=0, atol=1e-15)


def test_pf_reduce_and_S_fp64():
    m, o, g, rng = _setup(KIT3, 64, std=0.2)
    states = po.random_states(rng, 120, m.N)
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    got = g.reduce()
    assert np.allclose(got["o_mean"], ref["o_mean"], rtol=0, atol=1e-9 * (1 + np.abs(ref["o_mean"]).max()))
    assert np.allclose(got["eo"], ref["eo"], rtol=0, atol=1e-9 * (1 + np.abs(ref["eo"]).max()))
    x = rng.normal(size=m.P) + 1j * rng.normal(size=m.P)
    r1, r2, rd = 0.9, 9e-4, 1e-2
    y = g.apply_S(x, r1, r2, rd)
    # the Pfaffian block is regularised with r1 + rd (cg_solver.cpp:301-317)
    y_ref = po.apply_S(O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, x)
    assert np.max(np.abs(y - y_ref)) <= 1e-9 * np.abs(y_ref).max()


def test_pf_fp32_matches_exact_distribution():
    spec = ini("kitaev", 2, extra=PF + """
[sampler]
n_samples = 32768
metropolis.n_chains = 512
metropolis.n_warmup_steps = 50
pfaffian_refresh = 16
""")
    m, o, g, rng = _setup(spec, 32, std=0.3, seed=12)
    exact = o.exact_expectation()
    means = []
    for ep in range(4):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        std = r["std"]
    est = np.mean(means)
    assert abs(est - exact.real) < 5 * std / 2 + 1e-4, (est, exact, std)
