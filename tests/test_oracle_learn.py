"""Oracle pins for the learned proposal (the build's own design; the
reference has no learned proposal): the FP64 proposal network and its
batch gradient, and the cluster catalogue.

Pins: central finite differences of the loss for every parameter block, the
closed form of the loss at zero weights (uniform softmax), and the loops of
the lattice (hexagons / vertex stars, pinned by the lattice tests).
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import ini


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
