"""Shared fixtures: the model specs used across tests, as .ini text."""
import numpy as np

KITAEV_2x2 = """
[model]
type = kitaev
n_cells = 2
J = -1
symmetry = 0.5
[rbm]
type = basic
alpha = 1
"""

KITAEV_3x3 = KITAEV_2x2.replace("n_cells = 2", "n_cells = 3")


def ini(model="kitaev", n_cells=2, rbm="basic", alpha=1, symmetry=None, extra=""):
    sym = symmetry if symmetry is not None else (0.5 if model == "kitaev" else 1)
    return f"""
[model]
type = {model}
n_cells = {n_cells}
J = -1
symmetry = {sym}
[rbm]
type = {rbm}
alpha = {alpha}
{extra}
"""


def random_params(rng, P, std=0.1, std_imag=None):
    si = std if std_imag is None else std_imag
    return rng.normal(0, std, P) + 1j * rng.normal(0, si, P)


def learn_net(rng, N, Hn, s1=0.4, s2=0.3):
    """proposal network (W1, b1, W2, b2): W1 ~ s1 N(0, 1) (Hn x N), W2 ~ s2
    N(0, 1) (N x Hn), biases 0.1 N(0, 1), drawn in that order"""
    W1 = s1 * rng.normal(size=(Hn, N))
    b1 = 0.1 * rng.normal(size=Hn)
    W2 = s2 * rng.normal(size=(N, Hn))
    b2 = 0.1 * rng.normal(size=N)
    return W1, b1, W2, b2
