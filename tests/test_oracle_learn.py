"""Oracle pins for the learned proposal (the build's own design; the
reference has no learned proposal): the FP64 proposal network and its
batch gradient, and the cluster catalogue.

Pins: central finite differences of the loss for every parameter block, the
closed form of the loss at zero weights (uniform softmax), the loops of the
lattice (hexagons / vertex stars, pinned by the lattice tests), and detailed
balance of the exact transition matrix of the learned-only chain.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import ini, learn_net, random_params


def _net(rng, N, Hn, scale=0.3):
    return (scale * rng.normal(size=(Hn, N)), scale * rng.normal(size=Hn),
            scale * rng.normal(size=(N, Hn)), scale * rng.normal(size=N))


def test_mlp_logits_closed_form():
    rng = np.random.default_rng(0)
    N, Hn = 18, 16
    W1, b1, W2, b2 = _net(rng, N, Hn)
    st = po.random_states(rng, 5, N)
    z = po.mlp_logits(W1, b1, W2, b2, st, N)
    s = po.sigma_of(st, N)
    ref = np.maximum(s @ W1.T + b1, 0) @ W2.T + b2
    assert np.allclose(z, ref, rtol=1e-13, atol=1e-13)


def test_mlp_loss_at_zero_weights():
    """zero weights: p = 1/N, H = ln N, loss = mean(w) ln N - beta ln N"""
    N, Hn = 8, 4
    Z = (np.zeros((Hn, N)), np.zeros(Hn), np.zeros((N, Hn)), np.zeros(N))
    rng = np.random.default_rng(1)
    st = po.random_states(rng, 6, N)
    w = rng.normal(size=6)
    loss, g = po.mlp_grad(*Z, st, N, rng.integers(0, N, 6), w, 0.3)
    assert abs(loss - (w.mean() * np.log(N) - 0.3 * np.log(N))) < 1e-14
    # with W2 = 0 the hidden layer gets no gradient
    assert np.abs(g[0]).max() == 0 and np.abs(g[1]).max() == 0


@pytest.mark.parametrize("beta", [0.0, 0.05])
def test_mlp_grad_finite_differences(beta):
    rng = np.random.default_rng(2)
    N, Hn, n = 12, 10, 7
    net = list(_net(rng, N, Hn))
    st = po.random_states(rng, n, N)
    seeds = rng.integers(0, N, n)
    w = rng.normal(size=n)
    loss, g = po.mlp_grad(*net, st, N, seeds, w, beta)
    h = 1e-6
    for blk in range(4):
        flat = net[blk].reshape(-1)
        for idx in rng.choice(flat.size, min(flat.size, 12), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = po.mlp_grad(*net, st, N, seeds, w, beta)
            flat[idx] = old - h
            lm, _ = po.mlp_grad(*net, st, N, seeds, w, beta)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[blk].reshape(-1)[idx]) < 1e-7 * (1 + abs(fd)), (blk, idx, fd)


@pytest.mark.parametrize("model,n,kmax", [("kitaev", 2, 6), ("kitaev", 3, 6), ("toric", 3, 4)])
def test_cluster_catalogue(model, n, kmax):
    o = po.Oracle(ini(model, n))
    cyc = o.clusters()
    assert cyc.shape == (o.N, kmax)
    assert np.array_equal(cyc[:, 0], np.arange(o.N))
    loops = {tuple(sorted(h)) for h in o.hexagons()} if model == "kitaev" else \
        {tuple(sorted(p[:4])) for p in o.plaqs() if p[4] == 1}
    for i in range(o.N):
        assert tuple(sorted(cyc[i])) in loops
        assert len(set(cyc[i])) == kmax


def _learned_chain(o, net, N, stale_reverse=False):
    """exact transition matrix (2^N x 2^N) of the learned-only cluster chain
    (fraction 1) and |psi|^2: seed s ~ softmax(z(x)), cluster size k ~
    U{1..kmax}, flips = the first k sites of s's catalogue row; the forward
    and reverse proposal probabilities sum softmax(z(x)), softmax(z(y)) over
    every seed whose first k sites are the same set (the sampler's
    seed_mass); stale_reverse=True plants the mistake of taking the reverse
    one from the current state's logits z(x)"""
    cyc = np.asarray(o.clusters())
    kmax = cyc.shape[1]
    S = np.arange(1 << N, dtype=np.uint64).reshape(-1, 1)
    lp = o.logpsi(S)
    pi = np.exp(2 * (lp.real - lp.real.max()))
    pi /= pi.sum()
    z = po.mlp_logits(*net, S, N)
    q = np.exp(z - z.max(1, keepdims=True))
    q /= q.sum(1, keepdims=True)
    T = np.zeros((1 << N, 1 << N))
    for s in range(N):
        for k in range(1, kmax + 1):
            cl = set(cyc[s, :k].tolist())
            seeds = [j for j in range(N) if set(cyc[j, :k].tolist()) == cl]
            m = sum(1 << f for f in cl)
            for x in range(1 << N):
                y = x ^ m
                qf, qr = q[x, seeds].sum(), q[x if stale_reverse else y, seeds].sum()
                a = min(1.0, pi[y] * qr / (pi[x] * qf))
                T[x, y] += q[x, s] / kmax * a
                T[x, x] += q[x, s] / kmax * (1 - a)
    return T, pi


@pytest.mark.parametrize("scale", [0.3, 0.8])
def test_learned_chain_exact_balance_and_mixing(scale):
    """Kitaev 2x2 (N = 8, 6-site hexagon clusters, so clusters of size 6
    have 6 seeds), the networks of the GPU sampler test: the chain with the
    cluster-mass proposal probabilities satisfies detailed balance exactly
    (pi_x T_xy = pi_y T_yx); the planted variant that takes the reverse
    proposal probability from the current state's logits violates it.  Mixing: |lambda_2| of the 0.3-scale
    network's chain is below 0.99 (relaxation ~70 slots); that of the
    0.8-scale one is above 1 - 1e-5, a chain that no finite run equilibrates"""
    N = 8
    o = po.Oracle(ini("kitaev", 2, extra="[learn]\nhidden = 32\nfraction = 1.0\n"))
    rng = np.random.default_rng(5)
    o.set_params(random_params(rng, len(o.get_params()), 0.4))
    net = learn_net(rng, N, 32, scale, scale)
    T, pi = _learned_chain(o, net, N)
    assert np.allclose(T.sum(1), 1, atol=1e-13)
    F = pi[:, None] * T
    assert np.abs(F - F.T).max() < 1e-15
    assert np.abs(pi @ T - pi).max() < 1e-14
    lam2 = np.sort(np.abs(np.linalg.eigvals(T)))[::-1][1]
    if scale == 0.3:
        assert lam2 < 0.99, lam2
    else:
        assert lam2 > 1 - 1e-5, lam2
    Tb, _ = _learned_chain(o, net, N, stale_reverse=True)
    Fb = pi[:, None] * Tb
    assert np.abs(Fb - Fb.T).max() > 1e-4
