"""Host optimizer arithmetic (rbm-nqs_amd/host/optim.cpp through the host C
ABI) against the oracle's numpy restatement of plugin.cpp:30-77, and the
restatement itself against closed forms:
  * Adam, constant step g: m_hat = g and w_hat = g^2 exactly after the bias
    corrections, so every update is lr (Re g / (|Re g| + eps) + i Im g /
    (|Im g| + eps)) -- one step suffices to see a dropped correction;
  * momentum with dialup = 1: the first call starts from m = g, so a
    constant g gives lr g forever; with dialup the mixing factor is
    alpha * 1e-3 * dialup^t on call t.
"""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu


def _steps(rng, n, k):
    return [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(k)]


def test_adam_oracle_closed_form():
    g = np.array([0.3 - 2j, -1e-3 + 4j, 5.0 + 0j])
    for out in po.adam_steps([g] * 5, lr=0.01, eps=1e-8):
        want = 0.01 * (g.real / (np.abs(g.real) + 1e-8) + 1j * g.imag / (np.abs(g.imag) + 1e-8))
        assert np.allclose(out, want, rtol=1e-12, atol=1e-15)


def test_momentum_oracle_closed_form():
    g = np.array([0.3 - 2j, 1 + 1j])
    for out in po.momentum_steps([g] * 4, lr=0.1, alpha=0.9):
        assert np.allclose(out, 0.1 * g, rtol=1e-14)
    g2 = np.array([1.0 + 0j, -2j])
    out = po.momentum_steps([g, g2], lr=1.0, alpha=0.9, dialup=2.0)
    # call 1: m = g, a = 0.9 * 2e-3; call 2: a = 0.9 * 4e-3
    a2 = 0.9 * 4e-3
    assert np.allclose(out[1], a2 * g + (1 - a2) * g2, rtol=1e-14)


@pytest.mark.parametrize("kind,args", [("adam", (0.9, 0.999, 1e-8)), ("adam", (0.5, 0.9, 1e-3)),
                                       ("momentum", (0.9, 1.0, 0.0)), ("momentum", (0.7, 1.5, 0.0))])
def test_plugins_match_oracle(kind, args):
    rng = np.random.default_rng(2)
    dws = _steps(rng, 257, 6)
    pl = rbmgpu.Plugin(kind, 257, *args)
    got = [pl.apply(d, 0.03) for d in dws]
    if kind == "adam":
        ref = po.adam_steps(dws, 0.03, *args)
    else:
        ref = po.momentum_steps(dws, 0.03, alpha=args[0], dialup=args[1])
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-13, atol=1e-16)


def test_unknown_plugin_is_refused():
    with pytest.raises(RuntimeError, match="unknown plugin"):
        rbmgpu.Plugin("heun2", 4)
