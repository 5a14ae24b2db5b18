"""GPU parity tests: the HIP path (through the C ABI) against the FP64 CPU
oracle on the same inputs.

Tolerances: FP64 mode 1e-10 relative; FP32 mode as stated
per test (FP32 arithmetic, FP64 accumulation of observables).  Imaginary
parts of log psi are compared modulo 2 pi (sum of principal arguments).
FP32 ratios carry their rounding in the log-ratio, so their relative
tolerance is 2e-5 * (1 + |ln|ratio||).
"""
import sys

import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, random_params

pytestmark = pytest.mark.gpu

C1 = ini("kitaev", 3, extra="[sampler]\nn_samples = 1000\nmetropolis.n_chains = 4\n")
C2_SMALL = ini("toric", 4, alpha=2, extra="[sampler]\nn_samples = 128\nmetropolis.n_chains = 16\n")
C3_SMALL = ini("kitaev", 6, rbm="symmetry", alpha=4,
               extra="[sampler]\nn_samples = 64\nmetropolis.n_chains = 16\n")
SYM_TORIC = ini("toric", 3, rbm="symmetry", alpha=2, extra="[sampler]\nn_samples = 64\nmetropolis.n_chains = 8\n")
SPECS = {"C1": C1, "C2": C2_SMALL, "C3": C3_SMALL, "symtoric": SYM_TORIC}


def _wrap(x):
    return (x + np.pi) % (2 * np.pi) - np.pi


KEEP_O = "[gpu]\nkeep_o = true\n"


def _setup(spec, prec, std=0.3, seed=1, keep_o=False):
    print(f"[gpu-test] spec={spec.split()[3:6]} prec={prec} std={std} keep_o={keep_o}", flush=True)
    rng = np.random.default_rng(seed)
    m = rbmgpu.Model(spec + (KEEP_O if keep_o else ""))
    o = po.Oracle(spec)
    p = random_params(rng, m.P, std)
    o.set_params(p)
    g = m.gpu(precision=prec)
    g.set_params(p)
    return m, o, g, rng


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_logpsi(name, prec):
    m, o, g, rng = _setup(SPECS[name], prec)
    states = po.random_states(rng, 64, m.N)
    a = g.eval_logpsi(states)
    b = o.logpsi(states)
    tol = 1e-10 if prec == 64 else 2e-5
    scale = 1 + np.abs(b.real)
    assert np.all(np.abs(a.real - b.real) <= tol * scale), np.max(np.abs(a.real - b.real) / scale)
    assert np.all(np.abs(_wrap(a.imag - b.imag)) <= tol * scale)


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_psi_ratio(name, prec):
    m, o, g, rng = _setup(SPECS[name], prec)
    for _ in range(8):
        s = po.random_states(rng, 1, m.N)
        f = rng.choice(m.N, int(rng.integers(1, 7)), replace=False)
        a = g.psi_ratio(s, f)
        b = o.psi_ratio(s, f)
        tol = 1e-10 if prec == 64 else 2e-5 * (1 + abs(np.log(abs(b))))
        assert abs(a - b) <= tol * abs(b), (a, b)


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("prec", [64, 32])
def test_eloc_and_O(name, prec):
    m, o, g, rng = _setup(SPECS[name], prec, std=0.3 if prec == 64 else 0.1, keep_o=True)
    states = po.random_states(rng, 96, m.N)
    g.load_samples(states)
    st, obs, O, _ = g.get_samples(states=True, obs=True, O=True)
    assert np.array_equal(st, states)
    e_ref = o.eloc(states)
    O_ref = o.derivative(states)
    tol = 1e-10 if prec == 64 else 1e-4
    escale = np.abs(e_ref).max() + 1
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= tol * escale
    oscale = np.abs(O_ref).max()
    assert np.max(np.abs(O - O_ref)) <= tol * oscale


def test_structured_mode_has_no_O():
    """A basic RBM in FP32 runs the structured S.x: the O matrix is never
    formed, so asking for it is an error (rbmgpu.h keep_o)."""
    m, o, g, rng = _setup(C1, 32, std=0.1)
    g.load_samples(po.random_states(rng, 16, m.N))
    with pytest.raises(RuntimeError, match="structured"):
        g.get_samples(O=True)


# (spec, precision, keep_o): FP32 basic RBMs run the structured S.x unless
# keep_o; symmetric RBMs and FP64 always materialise O
REDUCE_CASES = [("C1", 64, False), ("C1", 32, False), ("C1", 32, True), ("C2", 64, False),
                ("C2", 32, False), ("C2", 32, True), ("C3", 64, False), ("C3", 32, False)]


# the CG loop queues its matvecs in batches, the first as long as the last
# solve: a run of solves whose iteration counts rise and fall crosses batch
# ends in every way (fixed counts), and a solve that converges mid-batch
# stops where the oracle's does (cg_solver.cpp:151-189)
@pytest.mark.gpu
def test_cg_batches_and_early_stop():
    m, o, g, rng = _setup(SPECS["C1"], 64, std=0.2)
    S = 200
    states = po.random_states(rng, S, m.N)
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    g.reduce()
    r1, r2, rd = 0.9, 9e-4, 1e-2
    F = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    x0 = 0.01 * (rng.normal(size=m.P) + 1j * rng.normal(size=m.P))
    args = (O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, F, x0)
    for mi in (2, 10, 3, 9, 1):
        xg, itg, rsg = g.sr_solve(F, x0, r1, r2, rd, max_iter=mi, rtol=1e-300)
        xo, ito, rso = po.cg(*args, mi, 1e-300)
        assert itg == ito == mi, mi
        assert np.max(np.abs(xg - xo)) <= 1e-9 * np.abs(xo).max(), mi
    # converge in the fifth step: rtol between the oracle's |r|^2 after 4
    # and 5 steps; the count is the loop index at the break, 4
    # (cg_solver.cpp:149-175)
    _, _, rs4 = po.cg(*args, 4, 1e-300)
    _, _, rs5 = po.cg(*args, 5, 1e-300)
    assert rs5 < rs4
    rtol = (rs4 * rs5) ** 0.5
    xo, ito, rso = po.cg(*args, 50, rtol)
    xg, itg, rsg = g.sr_solve(F, x0, r1, r2, rd, max_iter=50, rtol=rtol)
    assert ito == 4 and itg == ito
    assert np.max(np.abs(xg - xo)) <= 1e-9 * np.abs(xo).max()
    assert abs(rsg - rso) <= 1e-6 * rso


# energy moments at sample counts that give every range layout of the
# two-pass reduction: 50 bins of 120 (one range per bin), 32 bins of 2048
# (two ranges per bin), 2 bins of 1031 (a bin with no divisor in 64..1024:
# one range), a prime count (no binning, ragged last range)
@pytest.mark.gpu
@pytest.mark.parametrize("S", [6000, 65536, 2062, 3001])
def test_energy_moments_sample_counts(S):
    m, o, g, rng = _setup(SPECS["C1"], 64, std=0.2)
    states = po.random_states(rng, S, m.N)
    g.load_samples(states)
    ref = o.aggregate(o.eloc(states))
    got = g.reduce()
    tol = 1e-10
    assert abs(got["mean"] - ref["mean"]) <= tol * (1 + abs(ref["mean"]))
    for k in ("var", "var_binned", "tau", "std"):
        assert abs(got[k] - ref[k]) <= 10 * tol * (1 + abs(ref[k])), k


# every tile shape of the structured S.x: (model, n_cells, alpha) -> (CT, RT)
# = hidden / spin tiles of 16; (8, x) loads its chunks into registers, the
# others through the LDS chunk buffers
STRUCT_SHAPES = [("kitaev", 2, 1, "1x1"), ("kitaev", 2, 4, "2x1"), ("kitaev", 3, 1, "2x2"),
                 ("kitaev", 2, 8, "4x1"), ("toric", 4, 2, "4x2"), ("toric", 5, 1, "4x4"),
                 ("kitaev", 2, 16, "8x1"), ("kitaev", 4, 4, "8x2")]


@pytest.mark.parametrize("model,n,alpha,shape", STRUCT_SHAPES)
@pytest.mark.parametrize("S", [200, 5])
def test_structured_shapes(model, n, alpha, shape, S):
    """bf16x3 structured S.x and moments at every tile shape, for a sample
    count that is not a multiple of the 16-sample chunk and one smaller
    than a chunk.  The split products are exact, so the FP32 error comes
    from the FP32 tanh rows (measured <= 1.1e-4 of max at toric 5x5);
    stated tolerance 5e-4 of max, as for the materialised FP32 path."""
    spec = ini(model, n, alpha=alpha, extra="[sampler]\nn_samples = 200\nmetropolis.n_chains = 8\n")
    m, o, g, rng = _setup(spec, 32, std=0.1, seed=3)
    assert g.structured(), shape
    states = po.random_states(rng, S, m.N)
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    got = g.reduce()
    tol = 5e-4
    for k in ("o_mean", "eo"):
        assert np.max(np.abs(got[k] - ref[k])) <= tol * (1 + np.abs(ref[k]).max()), k
    assert np.max(np.abs(got["o_var"] - ref["o_var"])) <= tol * (1 + np.abs(ref["o_var"]).max())
    x = rng.normal(size=m.P) + 1j * rng.normal(size=m.P)
    r1, r2, rd = 0.9, 9e-4, 1e-2
    y = g.apply_S(x, r1, r2, rd)
    y_ref = po.apply_S(O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, x)
    assert np.max(np.abs(y - y_ref)) <= tol * np.abs(y_ref).max(), np.max(np.abs(y - y_ref)) / np.abs(y_ref).max()


@pytest.mark.parametrize("name,prec,keep_o", REDUCE_CASES)
def test_reduce_S_and_cg(name, prec, keep_o):
    m, o, g, rng = _setup(SPECS[name], prec, std=0.2 if prec == 64 else 0.1, keep_o=keep_o)
    S = 200
    states = po.random_states(rng, S, m.N)
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    got = g.reduce()
    tol = 1e-10 if prec == 64 else 1e-4
    assert abs(got["mean"] - ref["mean"]) <= tol * (1 + abs(ref["mean"]))
    for k in ("var", "var_binned", "tau"):
        assert abs(got[k] - ref[k]) <= 10 * tol * (1 + abs(ref[k])), k
    assert np.allclose(got["o_mean"], ref["o_mean"], rtol=0, atol=tol * (1 + np.abs(ref["o_mean"]).max()))
    assert np.allclose(got["o_var"], ref["o_var"], rtol=10 * tol, atol=10 * tol)
    assert np.allclose(got["eo"], ref["eo"], rtol=0, atol=tol * (1 + np.abs(ref["eo"]).max()))
    # S.x on the same sample set (cg_solver.cpp:301-317)
    x = rng.normal(size=m.P) + 1j * rng.normal(size=m.P)
    r1, r2, rd = 0.9, 9e-4, 1e-2
    y = g.apply_S(x, r1, r2, rd)
    y_ref = po.apply_S(O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, x)
    # FP32: O entries carry ~1e-6 relative error, amplified by cancellation
    # in the P-term products O_k . x (materialised) or the N- and M-term
    # sums of the structured form -> stated tolerance 5e-4 of max |y|
    stol = tol if prec == 64 else 5e-4
    assert np.max(np.abs(y - y_ref)) <= stol * np.abs(y_ref).max()
    # CG with a fixed iteration count and warm start (cg_solver.cpp:119-189);
    # 8 iterations stay far from convergence, where CG recurrences are
    # well conditioned (past convergence both sides divide rounding noise)
    F = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    x0 = 0.01 * (rng.normal(size=m.P) + 1j * rng.normal(size=m.P))
    xg, itg, rsg = g.sr_solve(F, x0, r1, r2, rd, max_iter=8, rtol=1e-300)
    xo, ito, rso = po.cg(O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, F, x0, 8, 1e-300)
    assert itg == ito == 8
    if prec == 64:
        assert np.max(np.abs(xg - xo)) <= 1e-9 * np.abs(xo).max()
        assert abs(rsg - rso) <= 1e-6 * rso
    else:
        # FP32 matvecs perturb the Krylov iterates of ill-conditioned random
        # sample sets: the FP64-evaluated residual of the GPU iterate is
        # within 5% of |F| of the oracle's residual (a side check; the
        # iterates themselves are compared in test_cg_fp32_iterates)
        res = lambda xx: np.linalg.norm(po.apply_S(O, ref["o_mean"], ref["o_var"], r1, r2, rd,
                                                   m.n_neural, xx) - F)
        assert res(xg) <= res(xo) + 0.05 * np.linalg.norm(F)


@pytest.mark.parametrize("name,keep_o", [("C1", False), ("C1", True), ("C2", False), ("C2", True),
                                         ("C3", False)])
def test_cg_fp32_iterates(name, keep_o):
    """FP32 S.x inside the FP64 CG: the iterate after 25 iterations against
    the oracle's FP64 CG on the same sample set.  A well-conditioned system
    (2000 samples, r2 = 0.1, so the shift r2 * max Var(O) bounds the
    condition number) keeps the Krylov recurrence from amplifying the
    ~1e-6 relative error of the FP32 products; stated bound: |x_gpu - x_o|
    <= 1e-3 |x_o| (max norm), and the iteration counts are equal."""
    m, o, g, rng = _setup(SPECS[name], 32, std=0.1, keep_o=keep_o)
    states = po.random_states(rng, 2000, m.N)
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    got = g.reduce()
    F = got["eo"] - got["mean"] * got["o_mean"].conj()
    F_ref = ref["eo"] - ref["mean"] * ref["o_mean"].conj()
    assert np.max(np.abs(F - F_ref)) <= 1e-4 * np.abs(F_ref).max()
    r1, r2, rd = 0.9, 0.1, 1e-2
    x0 = np.zeros(m.P, np.complex128)
    xg, itg, _ = g.sr_solve(F_ref, x0, r1, r2, rd, max_iter=25, rtol=1e-300)
    xo, ito, _ = po.cg(O, ref["o_mean"], ref["o_var"], r1, r2, rd, m.n_neural, F_ref, x0, 25, 1e-300)
    assert itg == ito == 25
    err = np.max(np.abs(xg - xo)) / np.abs(xo).max()
    print(f"[gpu-test] FP32 CG iterate error {err:.2e}", flush=True)
    assert err <= 1e-3, err


# Var(O) where |<O>|^2 >> Var(O): hidden units saturated by a large bias and
# spins frozen across the sample set, so tanh(theta_j) and sigma_i tanh
# theta_j are nearly constant.  The one-pass form (sum |O|^2 - S |<O>|^2) /
# (S - 1) loses every digit there; the reference accumulates Welford updates
# (aggregator.cpp:190-231) and so does the oracle.
VAR_CASES = [("C2", 64, False), ("C2", 32, False), ("C2", 32, True), ("C1", 32, False), ("C3", 64, False),
             ("C3", 32, False)]


@pytest.mark.parametrize("name,prec,keep_o", VAR_CASES)
def test_var_saturated_units(name, prec, keep_o):
    m, o, g, rng = _setup(SPECS[name], prec, std=0.01, keep_o=keep_o)
    p = o.get_params()
    p[m.n_vb:m.n_vb + m.n_alpha] = 3.0 + 0.2j  # hidden biases: tanh(theta) ~ 0.995
    o.set_params(p)
    g.set_params(p)
    S = 3000
    states = po.random_states(rng, S, m.N)
    states[:, 0] &= ~np.uint64(0xFF)  # spins 0..7 frozen at +1
    g.load_samples(states)
    e = o.eloc(states)
    O = o.derivative(states)
    ref = o.aggregate(e, O)
    got = g.reduce()
    zero = ref["o_var"] == 0  # the frozen spins' own entries: constant O
    assert np.all(got["o_var"][zero] == 0), np.max(got["o_var"][zero])
    ratio = np.abs(ref["o_mean"]) ** 2 / np.where(zero, 1.0, ref["o_var"])
    sat = (ratio >= 1e6) & ~zero
    assert sat.sum() >= m.n_alpha, "the sample set has no saturated parameters"
    # FP64: shifted deviations, 1e-10.  FP32: the FP32 tanh rows carry
    # ~6e-8 absolute rounding against deviations of ~1e-4..1e-3, so the
    # stated bound is 2e-3 relative on the saturated entries.
    tol = 1e-10 if prec == 64 else 2e-3
    rel = np.abs(got["o_var"][sat] - ref["o_var"][sat]) / ref["o_var"][sat]
    print(f"[gpu-test] {sat.sum()} saturated entries, max rel error {rel.max():.2e}", flush=True)
    assert rel.max() <= tol, rel.max()
    free = ~sat & ~zero
    ftol = 1e-10 if prec == 64 else 1e-4
    assert np.allclose(got["o_var"][free], ref["o_var"][free], rtol=10 * ftol, atol=10 * ftol)


def test_eloc_grid_stride():
    """k_eloc with more samples than one pass of its resident blocks covers
    (here 65536 kept samples of toric 4x4, alpha = 2: 16 ratio terms per
    sample, 32 samples per block, 3 blocks per CU x 256 CUs per pass), so
    every block walks the grid-stride loop several times; a spread subset
    of the kept samples is checked against the oracle."""
    spec = ini("toric", 4, alpha=2, extra="[sampler]\nn_samples = 65536\nmetropolis.n_chains = 4096\n"
                                          "metropolis.n_warmup_steps = 10\n")
    m, o, g, rng = _setup(spec, 32, std=0.1)
    g.sample_epoch(epoch=0)
    st, obs, _, _ = g.get_samples(states=True, obs=True, O=False)
    assert st.shape[0] == 65536
    idx = np.unique(np.concatenate([np.arange(0, 65536, 13), np.arange(65536 - 512, 65536)]))
    e_ref = o.eloc(st[idx])
    assert np.max(np.abs(obs[idx, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)


@pytest.mark.parametrize("name", ["C1", "C2", "symtoric", "C3"])
def test_sampler_trajectory_matches_oracle_fp64(name):
    """Same Philox streams and proposal rules: in FP64 the GPU chains follow
    the oracle's chains state for state."""
    m, o, g, rng = _setup(SPECS[name], 64, std=0.05)
    g.sample_epoch(epoch=3)
    st, obs, _, acc = g.get_samples(states=True, obs=True, O=False, chain_acc=True)
    S = m.n_samples
    st_o, e_o, _, acc_o = o.sample(seed=421390484, epoch=3, n_samples=S, n_chains=m.n_chains)
    same = np.all(st == st_o, axis=1)
    assert same.mean() == 1.0, f"{(~same).sum()} of {S} samples differ"
    assert np.max(np.abs(obs[:, 0] - e_o)) <= 1e-10 * (np.abs(e_o).max() + 1)
    assert np.allclose(acc, acc_o, rtol=0, atol=1e-15)


# the kept samples' local observables of an FP32 sampling epoch against the
# oracle's on the same states: basic RBMs with terms of at most 4 sites go
# through k_eloc after the sampler, the others through the sampler's own
# evaluation (local_op.cpp:31-65)
@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SPECS))
def test_sampler_fp32_local_energies(name):
    m, o, g, rng = _setup(SPECS[name], 32, std=0.1)
    g.sample_epoch(epoch=0)
    st, obs, _, _ = g.get_samples(states=True, obs=True, O=False)
    e_ref = o.eloc(st)
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)


def test_sampler_fp32_matches_exact_distribution():
    spec = ini("kitaev", 2, extra="""
[sampler]
n_samples = 65536
metropolis.n_chains = 256
metropolis.n_warmup_steps = 50
""")
    m, o, g, rng = _setup(spec, 32, std=0.4, seed=12)
    exact = o.exact_expectation()
    means = []
    for ep in range(4):
        g.sample_epoch(epoch=ep)
        r = g.reduce(want_o=False)
        means.append(r["mean"].real)
        std = r["std"]
    est = np.mean(means)
    assert abs(est - exact.real) < 5 * std / 2 + 1e-4, (est, exact, std)


def test_epoch_fp64_matches_oracle_epoch():
    """sample -> reduce -> F -> CG -> update, GPU vs oracle (SR.cpp:81-116)."""
    spec = ini("toric", 3, alpha=1, extra="""
[sampler]
n_samples = 400
metropolis.n_chains = 8
[optimizer]
sr.max_iterations = 50
""")
    m, o, g, rng = _setup(spec, 64, std=0.05)
    p0 = o.get_params()
    st = o.epoch(seed=421390484, epoch=0)  # oracle: lr=1e-2, reg1=0.9, reg2=9e-4 at epoch 1
    g.sample_epoch(epoch=0)
    r = g.reduce()
    F = r["eo"] - r["mean"] * r["o_mean"].conj()
    dw, itn, rs = g.sr_solve(F, np.zeros(m.P), 0.9, 9e-4, 1e-2, max_iter=50)
    assert itn == st["itn"]
    assert abs(r["mean"].real / m.N - st["energy"]) < 1e-12
    g.update_weights(1e-2 * dw)
    p_gpu = g.get_params()
    p_or = o.get_params()
    # 50 CG iterations amplify FP64 rounding differences to ~1e-8 relative
    assert np.max(np.abs(p_gpu - p_or)) <= 1e-7 * np.abs(p_or - p0).max()


@pytest.mark.parametrize("model,n,alpha,lanes", [("kitaev", 2, 4, 16), ("kitaev", 2, 4, 8), ("toric", 3, 2, 32)])
def test_group_sampler_symmetric_local_energies(model, n, alpha, lanes):
    """the lane-group FP32 sampler's batched local energies on symmetric
    RBMs (each unit reads the table row P[s][f] of its symmetry image s)
    against the oracle on its kept states, 1e-4 of the energy scale:
    Kitaev 2x2 alpha 4 (16 units; 16 lanes = one unit per lane, 8 lanes =
    two) and toric 3x3 alpha 2 (18 units on 32 lanes, 4-site terms)"""
    spec = ini(model, n, rbm="symmetry", alpha=alpha,
               extra="[sampler]\nn_samples = 512\nmetropolis.n_chains = 32\nmetropolis.n_warmup_steps = 10\n")
    print(f"[gpu-test] group sampler symmetric {model}{n} alpha={alpha} lanes={lanes}", flush=True)
    rng = np.random.default_rng(8)
    m = rbmgpu.Model(spec)
    o = po.Oracle(spec)
    p = random_params(rng, m.P, 0.1)
    o.set_params(p)
    g = m.gpu(precision=32, lanes=lanes)
    g.set_params(p)
    g.sample_epoch(epoch=0)
    assert not g.wave_sampler() and g.launch_shape()[0] == lanes, g.launch_shape()
    st, obs, _, _ = g.get_samples(states=True, obs=True, O=False)
    e_ref = o.eloc(st)
    assert np.max(np.abs(obs[:, 0] - e_ref)) <= 1e-4 * (np.abs(e_ref).max() + 1)
    assert len({tuple(r) for r in st}) > m.n_chains
