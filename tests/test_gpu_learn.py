"""GPU tests of the learned cluster proposal (no counterpart in the
reference; its pins are the FP64 oracle network and gradient, which are
checked against finite differences in test_oracle_learn.py, and exact
enumeration for the sampled distribution).

Tolerances: the training forward/backward runs in FP32 -> gradients within
1e-4 of the largest gradient entry; the sampler's forward of small
networks (hidden width and sites at most 64) runs FP32 on the bf16 weights
-> logits within 1e-4 of the logit scale, that of larger ones bf16 MFMA
with bf16 hidden activations -> within 2e-2, both against the oracle
evaluated on the same bf16-rounded weights.
"""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, learn_net, random_params

pytestmark = pytest.mark.gpu


def learn_spec(model, n, hidden, fraction, chains, samples, alpha=1, rbm="basic", warmup=100):
    return ini(model, n, rbm=rbm, alpha=alpha, extra=f"""
[sampler]
n_samples = {samples}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
[learn]
hidden = {hidden}
fraction = {fraction}
""")


_net = learn_net


def _flat(W1, b1, W2, b2):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2])


def _bf16(x):
    """round to bfloat16 (nearest even), returned as float64"""
    f = np.asarray(x, dtype=np.float32).copy()
    b = f.view(np.uint32)
    b += np.uint32(0x7FFF) + ((b >> np.uint32(16)) & np.uint32(1))
    b &= np.uint32(0xFFFF0000)
    return b.view(np.float32).astype(np.float64)


def _gpu(spec, seed=0, std=0.2):
    print(f"[gpu-test learn] {spec.split()[3:6]}", flush=True)
    m = rbmgpu.Model(spec)
    g = m.gpu(precision=32)
    rng = np.random.default_rng(seed)
    g.set_params(random_params(rng, m.P, std))
    return m, g, rng


@pytest.mark.parametrize("beta,n", [(0.0, 40), (0.05, 40), (0.05, 5000)])
def test_learn_grad_matches_oracle(beta, n):
    """the FP32 batch gradient against the FP64 oracle; 5000 records are 5
    batch slices of 1000 summed in slice order"""
    m, g, rng = _gpu(learn_spec("kitaev", 3, 48, 0.5, 4, 64))
    N, Hn = m.N, 48
    net = _net(rng, N, Hn)
    g.learn_weights(_flat(*net))
    net32 = [np.asarray(x, dtype=np.float32).astype(np.float64) for x in net]
    st = po.random_states(rng, n, N)
    seeds = rng.integers(0, N, n)
    seeds[3] = -1  # a record without a learned proposal contributes nothing
    w = rng.normal(size=n)
    loss, gg = g.learn_grad(st, seeds, w, beta)
    keep = seeds >= 0
    # the GPU averages over all n records; the oracle over the ones it is given
    lo, go = po.mlp_grad(*net32, st[keep], N, seeds[keep], w[keep], beta)
    scale = keep.sum() / n
    assert abs(loss - lo * scale) <= 1e-4 * (1 + abs(lo))
    ref = np.concatenate([x.ravel() for x in go]) * scale
    assert np.max(np.abs(gg - ref)) <= 1e-4 * np.abs(ref).max(), np.max(np.abs(gg - ref)) / np.abs(ref).max()


@pytest.mark.parametrize("Hn", [64, 80])
def test_learn_logits_matches_oracle(Hn):
    """the sampler's forward against the FP64 network on bf16 weights: at
    hidden width 64 (sites and hidden units at most 64) the per-wavefront
    FP32 forward, at 80 the bf16 MFMA one (operand layouts of
    v_mfma_f32_16x16x16_bf16)"""
    m, g, rng = _gpu(learn_spec("kitaev", 3, Hn, 0.5, 4, 64))
    N = m.N
    net = _net(rng, N, Hn, 0.3, 0.2)
    g.learn_weights(_flat(*net))
    st = po.random_states(rng, 21, N)
    z = g.learn_logits(st)
    W1, b1, W2, b2 = net
    zo = po.mlp_logits(_bf16(W1), np.float32(b1), _bf16(W2), np.float32(b2), st, N)
    tol = 1e-4 if Hn <= 64 else 2e-2
    assert np.max(np.abs(z - zo)) <= tol * (1 + np.abs(zo).max()), np.max(np.abs(z - zo))


@pytest.mark.parametrize("fraction,scale,chains", [(0.6, 0.8, 256), (1.0, 0.3, 256), (0.6, 0.8, 4096),
                                                   (0.6, 0.8, 518)])
def test_learned_sampler_matches_exact_distribution(fraction, scale, chains):
    """Metropolis-Hastings with the learned cluster proposal (untrained,
    non-uniform network) samples |psi|^2 exactly: <E> against exact
    enumeration within 5 binned standard errors, and the histogram of the
    sampled states within total variation 0.04 of |psi|^2.

    Learned moves alone (fraction 1) need a network whose chain mixes: with
    the 0.8-scale network the exact transition matrix of the learned-only
    chain has |lambda_2| = 1 - 2e-6 (relaxation ~4e5 slots), so no run of
    this length reaches |psi|^2 on any device (the exact finite-run
    distribution is 0.65 from it in total variation), while the 0.3-scale
    one relaxes in ~70 slots (test_learned_chain_exact_balance_and_mixing).
    At fraction 0.6 the plain single flips mix the 0.8-scale chain.  Blocks
    (this 8 -> 32 -> 8 network is evaluated per wavefront): 256 chains run
    one per block, 518 four per block (the last block holds two), 4096 are
    512 blocks of 8, more than the 256 CUs, so the launch takes the kernel
    variant limited to 128 VGPRs (two blocks per CU)."""
    spec = learn_spec("kitaev", 2, 32, fraction, chains, 32768, warmup=50)
    m, g, rng = _gpu(spec, seed=5, std=0.4)
    o = po.Oracle(spec)
    p = random_params(np.random.default_rng(5), m.P, 0.4)
    o.set_params(p)
    g.learn_weights(_flat(*_net(rng, m.N, 32, scale, scale)))
    exact = o.exact_expectation()
    lp = o.logpsi(np.arange(1 << m.N, dtype=np.uint64).reshape(-1, 1))
    pi = np.exp(2 * (lp.real - lp.real.max()))
    pi /= pi.sum()
    means, stds = [], []
    hist = np.zeros(1 << m.N)
    for ep in range(4):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        stds.append(r["std"])
        st = g.get_samples(states=True, obs=False)[0]
        hist += np.bincount(st[:, 0].astype(np.int64), minlength=1 << m.N)
    est = np.mean(means)
    err = np.sqrt(np.mean(np.square(stds)) / len(stds))
    assert abs(est - exact.real) < 5 * err + 1e-4, (est, exact, err)
    tv = 0.5 * np.abs(hist / hist.sum() - pi).sum()
    assert tv < 0.04, tv


def test_learned_sampler_mfma_16_chain_blocks_exact_energy():
    """the MFMA forward in blocks of 16 chains (the full 16-row tile): a
    18 -> 512 -> 18 network keeps ~100 KB of LDS per block, so one block
    fits a CU, and 4096 chains (512 blocks of 8, more than the 256 CUs)
    run as 256 blocks of 16 in the 128-VGPR variant; <E> of Kitaev 3x3
    against exact enumeration (2^18 states) within 5 binned standard
    errors"""
    spec = learn_spec("kitaev", 3, 512, 0.6, 4096, 32768, warmup=50)
    m, g, rng = _gpu(spec, seed=11, std=0.3)
    o = po.Oracle(spec)
    o.set_params(random_params(np.random.default_rng(11), m.P, 0.3))
    g.learn_weights(_flat(*_net(rng, m.N, 512, 0.4, 0.03)))
    exact = o.exact_expectation()
    means, stds = [], []
    for ep in range(3):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        stds.append(r["std"])
    est = np.mean(means)
    err = np.sqrt(np.mean(np.square(stds)) / len(stds))
    print(f"[gpu-test] 16-chain MFMA blocks: <E> {est} +- {err} (exact {exact.real})", flush=True)
    assert abs(est - exact.real) < 5 * err + 1e-4, (est, exact, err)


def test_learned_sampler_symmetric_local_energies():
    """the learned sampler on a symmetric RBM (Kitaev 2x2, alpha 4): the
    local energies of its kept states against the oracle, 1e-4 of the
    energy scale"""
    spec = learn_spec("kitaev", 2, 32, 0.5, 64, 2048, alpha=4, rbm="symmetry", warmup=20)
    m, g, rng = _gpu(spec, seed=9, std=0.1)
    o = po.Oracle(spec)
    o.set_params(random_params(np.random.default_rng(9), m.P, 0.1))
    g.learn_weights(_flat(*_net(rng, m.N, 32)))
    g.sample_epoch(epoch=0)
    st, obs, _, _ = g.get_samples(states=True, obs=True, O=False)
    e_ref = o.eloc(st)
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)


def test_learned_sampler_trains():
    """sampling -> records -> REINFORCE/Adam steps: finite loss, weights
    move, learned moves are proposed and accepted"""
    spec = learn_spec("kitaev", 3, 64, 0.5, 256, 8192, warmup=20)
    m, g, rng = _gpu(spec, seed=7, std=0.1)
    w0 = g.learn_weights()
    stats = []
    for ep in range(3):
        g.sample_epoch(epoch=ep)
        stats.append(g.learn_train(lr=1e-2, beta=0.01, batch=4096))
    w1 = g.learn_weights()
    assert all(np.isfinite(s["loss"]) for s in stats)
    assert stats[-1]["proposed"] > 0 and stats[-1]["accepted"] > 0
    assert np.max(np.abs(w1 - w0)) > 1e-4


# ---- the C5 shape: Kitaev 6x6 (N = 72, padded to 80 logit rows), symmetric
# RBM alpha = 4, proposal network 72 -> 256 -> 72
def c5_spec(chains=512, samples=4096, fraction=0.5, warmup=50):
    return learn_spec("kitaev", 6, 256, fraction, chains, samples, alpha=4, rbm="symmetry", warmup=warmup)


def test_learn_logits_c5_bf16_mfma():
    """72 -> 256 -> 72 bf16 MFMA forward of the sampler against the FP64
    network on the same bf16-rounded weights (pins the padded Np = 80
    operand layout and the 16 hidden tiles of 16)"""
    m, g, rng = _gpu(c5_spec(), std=0.05)
    N, Hn = m.N, 256
    assert N == 72
    net = _net(rng, N, Hn, 0.15, 0.1)
    g.learn_weights(_flat(*net))
    st = po.random_states(rng, 37, N)
    z = g.learn_logits(st)
    W1, b1, W2, b2 = net
    zo = po.mlp_logits(_bf16(W1), np.float32(b1), _bf16(W2), np.float32(b2), st, N)
    err = np.max(np.abs(z - zo)) / (1 + np.abs(zo).max())
    print(f"[gpu-test] C5 logits error {err:.2e} of scale {np.abs(zo).max():.2f}", flush=True)
    assert err <= 2e-2, err


@pytest.mark.parametrize("beta", [0.0, 0.05])
def test_learn_grad_c5(beta):
    """FP32 training gradient at the C5 shape against the FP64 oracle"""
    m, g, rng = _gpu(c5_spec(), std=0.05)
    N, Hn = m.N, 256
    net = _net(rng, N, Hn, 0.15, 0.1)
    g.learn_weights(_flat(*net))
    net32 = [np.asarray(x, dtype=np.float32).astype(np.float64) for x in net]
    n = 96
    st = po.random_states(rng, n, N)
    seeds = rng.integers(0, N, n)
    w = rng.normal(size=n)
    loss, gg = g.learn_grad(st, seeds, w, beta)
    lo, go = po.mlp_grad(*net32, st, N, seeds, w, beta)
    assert abs(loss - lo) <= 1e-4 * (1 + abs(lo))
    ref = np.concatenate([x.ravel() for x in go])
    err = np.max(np.abs(gg - ref)) / np.abs(ref).max()
    print(f"[gpu-test] C5 gradient error {err:.2e}", flush=True)
    assert err <= 1e-4, err


def test_learned_and_plain_agree_c5():
    """At the C5 shape the learned-move chains (lambda = 0.5, trained for two
    epochs) and plain Metropolis chains estimate the same <E>: difference
    within 5 combined binned standard errors over three epochs each"""
    m, g, rng = _gpu(c5_spec(chains=2048, samples=16384), seed=11, std=0.05)
    res = {}
    for name, frac in (("learned", 0.5), ("plain", 0.0)):
        g.learn_set_fraction(frac)
        if frac > 0:
            for ep in range(2):
                g.sample_epoch(epoch=100 + ep)
                g.learn_train(lr=1e-3, beta=0.01, batch=4096)
        means, stds = [], []
        for ep in range(3):
            g.sample_epoch(epoch=ep + (0 if frac > 0 else 50))
            r = g.reduce(want_o=False)
            means.append(r["mean"].real)
            stds.append(r["std"])
        res[name] = (np.mean(means), np.sqrt(np.mean(np.square(stds)) / len(stds)))
    (el, sl), (ep_, sp) = res["learned"], res["plain"]
    print(f"[gpu-test] C5 E learned {el:.5f} +- {sl:.5f}, plain {ep_:.5f} +- {sp:.5f}", flush=True)
    assert abs(el - ep_) < 5 * np.hypot(sl, sp), res


@pytest.mark.parametrize("fraction", [0.5, 1.0])
def test_learned_symmetric_matches_exact_distribution(fraction):
    """symmetric RBM (alpha = 2, 8 symmetries) under the learned cluster
    proposal at Kitaev 2x2: <E> against exact enumeration within 5 binned
    standard errors.  Weights N(0, 0.15^2): the sum over 8 symmetries makes
    the symmetric amplitude far more peaked than a basic RBM of the same
    weight scale (at 0.4 plain Metropolis accepts < 1% and its chains do
    not leave their starting modes, the oracle's as well); at 0.15 plain
    moves accept about 24%.  At lambda = 1 only one learned move in kmax = 6
    is a single flip, so the chains get 400 warm-up sweeps."""
    spec = learn_spec("kitaev", 2, 32, fraction, 256, 32768, alpha=2, rbm="symmetry", warmup=400)
    m, g, rng = _gpu(spec, seed=9, std=0.15)
    o = po.Oracle(spec)
    p = random_params(np.random.default_rng(9), m.P, 0.15)
    o.set_params(p)
    g.learn_weights(_flat(*_net(rng, m.N, 32, 0.8, 0.8)))
    exact = o.exact_expectation()
    means, stds = [], []
    for ep in range(4):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        stds.append(r["std"])
    est = np.mean(means)
    err = np.sqrt(np.mean(np.square(stds)) / len(stds))
    assert abs(est - exact.real) < 5 * err + 1e-4, (est, exact, err)


def test_learned_beats_single_flips_on_converged_state():
    """The learned proposal against the plain single-flip Metropolis that it
    falls back to, on a state where single flips barely move: the
    reference's Kitaev example (2x2, symmetric RBM alpha = 4) trained to the
    ED energy (rbm-nqs_amd/examples/kitaev2_sym_ed.rbm), 256 chains x 256
    samples per epoch.  Plain single flips are accepted ~6e-4 of the time;
    the learned cluster moves (lambda = 0.5, single-flip fallback, 20
    training epochs on the GPU) mix the chains, so at equal sample counts the
    rms error of E/N against the exact expectation value of the same weights
    (GPU exact enumeration, FP64) is smaller.  Both runs are deterministic
    (Philox streams); the measured ratio is about 0.5, the bound 0.8."""
    import os
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    mf = rbmgpu.Model(bench.ed2_spec(0.0, full=True))
    params = bench.read_rbm(bench.ED2_CKPT, mf)
    gf = mf.gpu(precision=64)
    gf.set_params(params)
    gf.full_epoch()
    exact = gf.reduce(want_o=False)["mean"].real / mf.N
    assert abs(exact + 0.8660254037844461) < 2e-5  # the checkpoint is converged
    rms = {}
    for learn in (0, 32):
        print(f"[gpu-test] converged Kitaev 2x2, learn={learn}", flush=True)
        m = rbmgpu.Model(bench.ed2_spec(0.0, learn=learn, lam=0.5))
        g = m.gpu(precision=32)
        g.set_params(params)
        if learn:
            for e in range(20):
                g.sample_epoch(epoch=1000 + e)
                g.learn_train(lr=3e-3, beta=0.01, batch=m.n_samples)
        err = []
        for e in range(8):
            g.sample_epoch(epoch=e)
            err.append(g.reduce(want_o=False)["mean"].real / m.N - exact)
        rms[learn] = float(np.sqrt(np.mean(np.square(err))))
    print(f"[gpu-test] rms error: plain single flips {rms[0]}, learned {rms[32]}", flush=True)
    assert rms[32] < 0.8 * rms[0], rms
