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
