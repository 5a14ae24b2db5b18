"""The one-chain-per-wavefront FP32 sampler (k_sample_wave, forced with
lanes = -1) against the FP64 oracle: the local energies it evaluates for
its kept samples and the O rows built from its stored tanh(theta) rows
agree with the oracle's on the same states (FP32 tolerance 1e-4 relative,
as for the lane-group sampler); its chains sample |psi|^2 (energy against
exact enumeration, 5 sigma) with single flips, composed bond flips (basic
RBM) and the pair tables of the symmetric RBM; its acceptance rate agrees
with the lane-group sampler's on the same model.  Covers one and several
units per lane (up to the largest, 8), padding units (H not a multiple of
64), 4-site terms and a sample count that does not divide by the chains.
"""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, random_params

pytestmark = pytest.mark.gpu

KEEP = "[gpu]\nkeep_o = true\n"


def _smp(n, chains, warmup=20):
    return f"[sampler]\nn_samples = {n}\nmetropolis.n_chains = {chains}\nmetropolis.n_warmup_steps = {warmup}\n"


SPECS = {
    # basic RBM, H = 18 (one unit per lane, 46 padding lanes)
    "kitaev3_basic": ini("kitaev", 3, extra=_smp(512, 16)),
    # basic RBM, H = 90: two units per lane, the second one padded
    "kitaev3_basic_a5": ini("kitaev", 3, alpha=5, extra=_smp(512, 16)),
    # basic RBM, H = 64, 4-site star / plaquette terms
    "toric4_basic_a2": ini("toric", 4, alpha=2, extra=_smp(512, 16)),
    # symmetric RBM alpha = 4 on 72 sites: five units per lane, pair tables
    "kitaev6_sym_a4": ini("kitaev", 6, rbm="symmetry", alpha=4, extra=_smp(256, 32)),
    # symmetric RBM, 4-site terms
    "toric3_sym_a2": ini("toric", 3, rbm="symmetry", alpha=2, extra=_smp(512, 16)),
    # symmetric RBM alpha = 2 on 32 sites
    "kitaev4_sym_a2": ini("kitaev", 4, rbm="symmetry", alpha=2, extra=_smp(512, 16)),
    # the largest shape: 512 hidden units, eight units per lane
    "kitaev4_sym_a16": ini("kitaev", 4, rbm="symmetry", alpha=16, extra=_smp(256, 16)),
    # ragged: 1003 samples over 16 chains (the remainder goes to chain 0)
    "kitaev3_ragged": ini("kitaev", 3, extra=_smp(1003, 16)),
}


def _gpu(spec, std, seed=1, lanes=-1, keep_o=True):
    rng = np.random.default_rng(seed)
    m = rbmgpu.Model(spec + (KEEP if keep_o else ""))
    o = po.Oracle(spec)
    p = random_params(rng, m.P, std)
    o.set_params(p)
    g = m.gpu(precision=32, lanes=lanes)
    g.set_params(p)
    return m, o, g


@pytest.mark.parametrize("name", list(SPECS))
def test_wave_sampler_local_energies_and_O(name):
    print(f"[gpu-test] wave sampler {name}", flush=True)
    # 512 hidden units at std 0.1 concentrate |psi|^2 on a few states
    m, o, g = _gpu(SPECS[name], 0.03 if name == "kitaev4_sym_a16" else 0.1)
    g.sample_epoch(epoch=0)
    assert g.wave_sampler(), g.launch_shape()
    st, obs, O, _ = g.get_samples(states=True, obs=True, O=True)
    assert st.shape[0] == m.n_samples
    e_ref = o.eloc(st)
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)
    O_ref = o.derivative(st)
    assert np.max(np.abs(O - O_ref)) <= 1e-4 * np.abs(O_ref).max()
    # the chains moved: more distinct states than chains
    assert len({tuple(r) for r in st}) > m.n_chains


DIST = {
    # basic RBM alpha = 2: composed bond flips, one unit per lane
    "basic": ini("kitaev", 2, alpha=2, extra=_smp(65536, 256, 50)),
    # symmetric RBM alpha = 16 (H = 128: two units per lane), pair tables
    "sym": ini("kitaev", 2, rbm="symmetry", alpha=16, extra=_smp(65536, 256, 50)),
    # single flips only (they mix slowly on the Kitaev model: smaller
    # weights and a longer warm-up)
    "single": ini("kitaev", 2, rbm="symmetry", alpha=4,
                  extra=_smp(65536, 256, 300) + "metropolis.bond_flips = 0\n"),
}
STD = {"basic": 0.4, "sym": 0.4, "single": 0.15}


@pytest.mark.parametrize("name", list(DIST))
def test_wave_sampler_matches_exact_distribution(name):
    print(f"[gpu-test] wave sampler distribution {name}", flush=True)
    m, o, g = _gpu(DIST[name], STD[name], seed=12, keep_o=False)
    exact = o.exact_expectation()
    means = []
    for ep in range(4):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        std = r["std"]
    assert g.wave_sampler()
    est = np.mean(means)
    print(f"[gpu-test] {name}: <E> {est} +- {std / 2} (exact {exact.real}), acceptance {r['acc']}", flush=True)
    assert abs(est - exact.real) < 5 * std / 2 + 1e-4, (est, exact, std)


def test_wave_and_group_samplers_accept_alike():
    spec = ini("kitaev", 6, rbm="symmetry", alpha=4, extra=_smp(16384, 2048, 20))
    accs = {}
    for lanes in (-1, 64):
        print(f"[gpu-test] acceptance lanes={lanes}", flush=True)
        m, o, g = _gpu(spec, 0.05, seed=3, lanes=lanes, keep_o=False)
        a = []
        for ep in range(2):
            g.sample_epoch(epoch=ep)
            a.append(g.reduce(want_o=False)["acc"])
        assert g.wave_sampler() == (lanes == -1)
        accs[lanes] = np.mean(a)
    # 2 x 2048 chains x 28 sweeps x ~48 proposals: binomial error ~1e-3
    assert 0.05 < accs[64] < 0.95
    assert abs(accs[-1] - accs[64]) < 0.01, accs


def test_auto_chain_count():
    """gpu.auto_chains: the autotuner picks the chain count (a power of two
    from 1024, or the sample count) before the first epoch; the epoch's
    kept samples keep their local-energy and O parity, every chain reports
    its acceptance"""
    spec = ini("kitaev", 6, rbm="symmetry", alpha=4, extra=_smp(65536, 64) + "[gpu]\nauto_chains = true\n")
    print("[gpu-test] auto chains", flush=True)
    rng = np.random.default_rng(4)
    m = rbmgpu.Model(spec + KEEP)
    o = po.Oracle(spec)
    p = random_params(rng, m.P, 0.05)
    o.set_params(p)
    g = m.gpu(precision=32)
    g.set_params(p)
    g.sample_epoch(epoch=0)
    n = g.chains()
    print(f"[gpu-test] picked {n} chains", flush=True)
    assert n in (1024, 2048, 4096, 8192, 16384, 32768, 65536)
    st, obs, O, ca = g.get_samples(states=True, obs=True, O=True, chain_acc=True)
    assert ca.shape == (n,) and np.all(ca > 0)
    sel = rng.choice(st.shape[0], 256, replace=False)
    e_ref = o.eloc(st[sel])
    assert np.max(np.abs(obs[sel, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)
    O_ref = o.derivative(st[sel])
    assert np.max(np.abs(O[:, sel] - O_ref)) <= 1e-4 * np.abs(O_ref).max()
