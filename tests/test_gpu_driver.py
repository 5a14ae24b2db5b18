"""End-to-end runs of the `rbm` driver on the GPU (main.cpp:381-717
semantics): training writes the <name>.log CSV (logger.cpp:31-74) and the
<name>.rbm checkpoint (rbm_base.cpp:219-288, eigen_fstream.cpp:25-51); a
second run resumes from the checkpoint's update counter; --evaluate prints
the reference's summary lines.  Energies are checked against the Kitaev
2x2 exact ground energy (ED, -0.8660254037844461 per site) only as bounds
(a few SR epochs from random weights).
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RBM = os.path.join(ROOT, "rbm-nqs_amd", "lib", "rbm")

INI = """
name = {name}
n_epochs = {epochs}
noprogress = true
[model]
type = kitaev
n_cells = 2
J = -1
symmetry = 0.5
[rbm]
type = basic
alpha = 1
weights = 0.01
[sampler]
n_samples = 2000
metropolis.n_chains = 100
[optimizer]
learning_rate = 0.05
"""


def _run(args, cwd):
    print("[gpu-test driver]", " ".join(args), flush=True)
    p = subprocess.run([RBM] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=180)
    return p.returncode, p.stdout


def _read_rbm(path, N, na):
    with open(path, "rb") as f:
        data = f.read()
    off, mats = 0, []
    for rows, cols in ((N, na), (na, 1), (N, 1)):
        r, c = struct.unpack_from("<QQ", data, off)
        assert (r, c) == (rows, cols)
        off += 16
        m = np.frombuffer(data, dtype=np.complex128, count=r * c, offset=off)
        mats.append(m.reshape(c, r).T)  # column-major
        off += 16 * r * c
    (nu,) = struct.unpack_from("<Q", data, off)
    return mats, nu


def test_driver_train_resume_evaluate(tmp_path):
    d = str(tmp_path)
    with open(os.path.join(d, "run.ini"), "w") as f:
        f.write(INI.format(name="k2", epochs=30))
    rc, out = _run(["-i", "run.ini", "--train"], d)
    assert rc == 0, out
    log = open(os.path.join(d, "k2.log")).read().splitlines()
    assert log[0] == "# Energy, Energy Stddev, AccetpanceRate"
    rows = np.array([[float(x) for x in l.split(",")] for l in log[1:]])
    assert rows.shape == (30, 3)
    # variational energies stay above the exact ground energy and drop
    assert np.all(rows[:, 0] > -0.8660254037844461 - 5 * rows[:, 1].max())
    assert rows[-5:, 0].mean() < rows[:5, 0].mean()
    (W, h, v), nu = _read_rbm(os.path.join(d, "k2.rbm"), 8, 8)
    assert nu == 30 and np.all(np.isfinite(W))
    # resume: n_epochs 40 continues from update 30 (10 more lines appended)
    with open(os.path.join(d, "run.ini"), "w") as f:
        f.write(INI.format(name="k2", epochs=40))
    rc, out = _run(["-i", "run.ini", "--train"], d)
    assert rc == 0, out
    assert "Loaded RBM from 'k2.rbm'!" in out
    log2 = open(os.path.join(d, "k2.log")).read().splitlines()
    assert len([l for l in log2 if not l.startswith("#")]) == 40
    _, nu2 = _read_rbm(os.path.join(d, "k2.rbm"), 8, 8)
    assert nu2 == 40
    # evaluate: hexagon plaquettes, variance, std, tau, acceptance, energy
    rc, out = _run(["-i", "run.ini", "--evaluate", "--sampler.eval_samples=4000"], d)
    assert rc == 0, out
    for key in ("Hex0:", "Var:", "Std:", "Tau:", "Acc:", "E:"):
        assert key in out, out
    e = float(re.search(r"E:\s+(\S+)", out).group(1))
    assert -0.8660254037844461 - 0.05 < e < 0


def test_driver_exact_energy_and_errors(tmp_path):
    rc, out = _run(["--exact_energy", "--model.n_cells=3"], str(tmp_path))
    assert rc == 0 and "Exact Energy: -0.79397236789" in out
    rc, out = _run(["--train", "--optimizer.sr.method=minresqlp"], str(tmp_path))
    assert rc != 0 and "minresqlp" in out


def _write_rbm(path, W, h, v, nu):
    with open(path, "wb") as f:
        for m in (W, h.reshape(-1, 1), v.reshape(-1, 1)):
            f.write(struct.pack("<QQ", *m.shape))
            f.write(np.asarray(m, np.complex128).T.tobytes())  # column-major
        f.write(struct.pack("<Q", nu))


# the optimizer family of main.cpp:241-287 (SR solvers, SGD, plugins,
# resample): each runs end to end and lowers the Kitaev 2x2 energy.  The
# arithmetic of each piece is pinned elsewhere (plugins against the
# oracle, explicit S / cg-single / direct against the oracle's S).
OPTIMIZERS = {
    "cg-single": "sr.method = cg-single\n",
    "direct": "sr.method = direct\n",
    "cg-dynamic": "sr.method = cg-dynamic\n",
    "minres": "sr.method = minres\n",
    "adam": "plugin = adam\n",
    # momentum accumulates the SR steps: with the default alpha 0.9 at this
    # learning rate the Kitaev 2x2 run reaches a step the guard refuses, so
    # the run here ramps the momentum up (dialup) from 1e-3
    "momentum": "plugin = momentum\nmom.dialup = 1.1\n",
    "heun": "plugin = heun\n",
    "sgd": "type = SGD\n",
    "resample": "resample = true\n",
}


@pytest.mark.parametrize("opt", list(OPTIMIZERS))
def test_driver_optimizer_family(tmp_path, opt):
    d = str(tmp_path)
    text = INI.format(name="o", epochs=60) + OPTIMIZERS[opt]
    if opt == "sgd":
        text = text.replace("learning_rate = 0.05", "learning_rate = 0.02")
    with open(os.path.join(d, "run.ini"), "w") as f:
        f.write(text)
    rc, out = _run(["-i", "run.ini", "--train"], d)
    assert rc == 0, out
    log = open(os.path.join(d, "o.log")).read().splitlines()
    rows = np.array([[float(x) for x in l.split(",")] for l in log[1:]])
    assert rows.shape[0] == 60 and np.all(np.isfinite(rows))
    assert np.all(rows[:, 0] > -0.8660254037844461 - 5 * rows[:, 1].max())
    assert rows[-10:, 0].mean() < rows[:10, 0].mean() - 0.01, (opt, rows[:10, 0].mean(), rows[-10:, 0].mean())


def test_driver_learned_proposal_trains(tmp_path):
    """`learn.hidden` switches on the learned cluster proposal in the driver:
    every epoch samples with learned moves (fraction 0.5) and runs one
    REINFORCE / Adam step of the network on the epoch's records; the SR
    run still descends (same bounds as the optimizer family)"""
    d = str(tmp_path)
    text = INI.format(name="l", epochs=60) + "[learn]\nhidden = 32\nfraction = 0.5\nbatch = 2000\n"
    with open(os.path.join(d, "run.ini"), "w") as f:
        f.write(text)
    rc, out = _run(["-i", "run.ini", "--train"], d)
    assert rc == 0, out
    log = open(os.path.join(d, "l.log")).read().splitlines()
    rows = np.array([[float(x) for x in l.split(",")] for l in log[1:]])
    assert rows.shape[0] == 60 and np.all(np.isfinite(rows))
    assert np.all(rows[:, 0] > -0.8660254037844461 - 5 * rows[:, 1].max())
    assert rows[-10:, 0].mean() < rows[:10, 0].mean() - 0.01, (rows[:10, 0].mean(), rows[-10:, 0].mean())


@pytest.mark.parametrize("extra", ["", "[optimizer]\nplugin = momentum\n"])
def test_driver_massive_step_exits_55(tmp_path, extra):
    """a checkpoint with a NaN weight: every local energy, F and step is NaN,
    the guard refuses the step (abstract_optimizer.cpp:157) and the driver
    exits with rc 55 (main.cpp:607-611) -- on the device epoch and on the
    host-driven path (a plugin)"""
    d = str(tmp_path)
    rng = np.random.default_rng(1)
    W = 0.01 * (rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))
    W[2, 3] = np.nan
    _write_rbm(os.path.join(d, "nan.rbm"), W, np.zeros(8, complex), np.zeros(8, complex), 0)
    with open(os.path.join(d, "run.ini"), "w") as f:
        f.write(INI.format(name="nan", epochs=5) + extra)
    rc, out = _run(["-i", "run.ini", "--train"], d)
    assert rc == 55, out
    assert "FATAL: Massive step detected." in out


EXAMPLES = os.path.join(ROOT, "rbm-nqs_amd", "examples")


def _energy(out):
    m = re.search(r"E:\s+(\S+) \xb1 (\S+)", out)
    return float(m.group(1)), float(m.group(2))


def test_driver_toric4_converges_to_reference(tmp_path):
    """examples/toric4.ini (the settings of the reference's toric_n4 run):
    1000 SR epochs then --evaluate.  The reference reached E/N = -0.9999703
    +- 2e-5 (examples/toric/backup/toric_n4.plog, OUTPUT block); the exact
    value is -1.  Ours must agree with the reference within the combined
    statistical error and lie within 1e-4 of the exact energy."""
    rc, out = _run(["-i", os.path.join(EXAMPLES, "toric4.ini"), "--train", "--evaluate"], str(tmp_path))
    assert rc == 0, out
    e, sd = _energy(out)
    print(f"[gpu-test] toric 4x4: E/N = {e} +- {sd}", flush=True)
    assert abs(e + 1) <= 1e-4
    assert abs(e - (-0.9999703)) <= 3 * np.hypot(2e-5, sd)


def test_driver_reference_kitaev_example_reaches_ed(tmp_path):
    """examples/kitaev2_full.ini is the reference's examples/kitaev/kitaev.ini
    (Kitaev 2x2, symmetric RBM alpha = 4, exact enumeration, SR with the
    direct solver, 3000 epochs).  The first 1300 epochs of that run take the
    energy per site to the exact-diagonalisation value -0.8660254037844461
    (notebooks/KitaevED.ipynb) within 1e-4, from above (exact enumeration:
    no statistical error)"""
    rc, out = _run(["-i", os.path.join(EXAMPLES, "kitaev2_full.ini"), "--train", "--evaluate", "--n_epochs=1300"],
                   str(tmp_path))
    assert rc == 0, out
    e, _ = _energy(out)
    print(f"[gpu-test] Kitaev 2x2 (reference example): E/N = {e}", flush=True)
    assert -0.8660254037844461 - 1e-9 <= e <= -0.8660254037844461 + 1e-4
