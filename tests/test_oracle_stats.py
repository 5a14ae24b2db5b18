"""Pins the oracle's aggregators, S.x, CG and Metropolis sampler with closed
forms, explicit matrices and exact enumeration."""
import numpy as np
import pytest

import pyoracle as po
from helpers import ini, random_params


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    assert po.philox([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert po.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert po.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]).tolist() == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_aggregator_closed_forms():
    # aggregator.cpp:84-239 for one rank with unit weights
    rng = np.random.default_rng(5)
    o = po.Oracle(ini("kitaev", 2))
    S = 1000
    e = rng.normal(size=S) + 1j * rng.normal(size=S) * 0.1
    e = np.cumsum(e) * 0.01 + e  # autocorrelated stream
    O = rng.normal(size=(o.P, S)) + 1j * rng.normal(size=(o.P, S))
    r = o.aggregate(e, O)
    assert abs(r["mean"] - e.mean()) < 1e-12
    var = np.sum(np.abs(e - e.mean()) ** 2) / (S - 1)
    assert abs(r["var"] - var) < 1e-10 * var
    nb = 50
    bs = S // nb
    bins = e.reshape(nb, bs).mean(axis=1)
    vb = bs * np.sum(np.abs(bins - bins.mean()) ** 2) / (S - 1)
    assert abs(r["var_binned"] - vb) < 1e-10 * vb
    assert abs(r["tau"] - 0.5 * (max(bs * vb / var, 1) - 1)) < 1e-9
    assert abs(r["std"] - np.sqrt(vb / nb)) < 1e-12
    assert np.allclose(r["o_mean"], O.mean(axis=1), atol=1e-13)
    ovar = np.sum(np.abs(O - O.mean(axis=1, keepdims=True)) ** 2, axis=1) / (S - 1)
    assert np.allclose(r["o_var"], ovar, rtol=1e-10)
    assert np.allclose(r["eo"], (e[None, :] * O.conj()).mean(axis=1), atol=1e-13)


def _explicit_S(O, mean, var, r1, r2, rd, nn):
    P, S = O.shape
    Sm = O.conj() @ O.T / S - np.outer(mean.conj(), mean)
    reg = var * np.where(np.arange(P) < nn, r1, r1 + rd) + var.max() * r2
    return Sm + np.diag(reg)


def test_apply_S_matches_explicit_matrix():
    # cg_solver.cpp:301-317 vs the explicit matrix of cg_solver.cpp:274-282
    rng = np.random.default_rng(6)
    P, S, nn = 40, 300, 30
    O = rng.normal(size=(P, S)) + 1j * rng.normal(size=(P, S))
    mean = O.mean(axis=1)
    var = np.sum(np.abs(O - mean[:, None]) ** 2, axis=1) / (S - 1)
    x = rng.normal(size=P) + 1j * rng.normal(size=P)
    y = po.apply_S(O, mean, var, 0.9, 9e-4, 1e-2, nn, x)
    Sm = _explicit_S(O, mean, var, 0.9, 9e-4, 1e-2, nn)
    assert np.allclose(y, Sm @ x, rtol=1e-12, atol=1e-12)
    assert np.allclose(Sm, Sm.conj().T, atol=1e-12)  # Hermitian


def test_cg_solves_system_with_warm_start_semantics():
    rng = np.random.default_rng(8)
    P, S = 30, 400
    O = rng.normal(size=(P, S)) + 1j * rng.normal(size=(P, S))
    mean = O.mean(axis=1)
    var = np.sum(np.abs(O - mean[:, None]) ** 2, axis=1) / (S - 1)
    b = rng.normal(size=P) + 1j * rng.normal(size=P)
    Sm = _explicit_S(O, mean, var, 0.5, 1e-3, 0.0, P)
    x, itn, rs = po.cg(O, mean, var, 0.5, 1e-3, 0.0, P, b, np.zeros(P), 2 * P, 1e-20)
    assert np.allclose(x, np.linalg.solve(Sm, b), rtol=1e-8, atol=1e-10)
    assert rs < 1e-20 and itn < 2 * P
    # warm start at the solution: converges at once
    x2, itn2, rs2 = po.cg(O, mean, var, 0.5, 1e-3, 0.0, P, b, x, 2 * P, 1e-16)
    assert itn2 == 0
    # iteration cap: itn == max_iter when rtol is unreachable
    _, itn3, _ = po.cg(O, mean, var, 0.5, 1e-3, 0.0, P, b, np.zeros(P), 3, 0.0)
    assert itn3 == 3


def test_sampler_matches_exact_enumeration():
    # stationary distribution of the Metropolis chains == |psi|^2
    rng = np.random.default_rng(12)
    spec = ini("kitaev", 2, extra="""
[sampler]
n_samples = 40000
metropolis.n_chains = 16
metropolis.n_warmup_steps = 50
metropolis.bond_flips = 0.5
""")
    o = po.Oracle(spec)
    o.set_params(random_params(rng, o.P, 0.4))
    exact = o.exact_expectation()
    _, e, _, acc = o.sample(seed=421390484, epoch=0, n_samples=40000, n_chains=16)
    agg = o.aggregate(e)
    assert 0.001 < acc.mean() < 1.0
    assert abs(agg["mean"].real - exact.real) < 5 * agg["std"] + 1e-3


def test_sampler_is_deterministic_per_chain_stream():
    rng = np.random.default_rng(13)
    spec = ini("toric", 2, extra="[sampler]\nn_samples = 64\nmetropolis.n_chains = 4\n")
    o = po.Oracle(spec)
    o.set_params(random_params(rng, o.P, 0.3))
    a = o.sample(1, 0, 64, 4, threads=1)
    b = o.sample(1, 0, 64, 4, threads=4)
    assert (a[0] == b[0]).all() and np.array_equal(a[1], b[1])
    c = o.sample(1, 1, 64, 4, threads=1)
    assert not (a[0] == c[0]).all()


def test_epoch_lowers_energy_kitaev():
    # SR epochs on the 2x2 Kitaev model move towards the ED energy -0.866
    spec = ini("kitaev", 2, extra="""
[sampler]
n_samples = 2000
metropolis.n_chains = 8
[optimizer]
learning_rate = 0.1
""")
    o = po.Oracle(spec)
    rng = np.random.default_rng(14)
    o.set_params(random_params(rng, o.P, 0.05))
    es = [o.epoch(seed=7, epoch=i)["energy"] for i in range(30)]
    assert es[0] > -0.55
    assert np.mean(es[-5:]) < -0.72
    assert np.mean(es[-5:]) > -0.8660254037844461 - 0.01  # variational bound
