"""Benchmark of the VMC epoch on one MI355X.

Headline workload (`value`): Kitaev honeycomb 6x6 (N = 72 spins), symmetric
RBM alpha = 4 with the 72 translation x rotation symmetries (P = 293),
16384 Metropolis chains x 8 kept samples (131 072 samples per epoch), 100
warm-up sweeps, bond_flips 0.5, FP32 sampler and O contractions, SR solved
by CG with FP64 vectors (sr.max_iterations = 100).  One step = one whole SR
epoch on the device (rbmgpu_train_epoch: sample -> local energies ->
moments -> F -> CG -> guard -> update), captured once in a HIP graph and
replayed.  Weights: the checkpoint rbm-nqs_amd/examples/kitaev6_sym.rbm
(trained by the `rbm` driver on kitaev6_sym.ini) when present, else
N(0, 0.05^2) from the reference's initialiser.

`value` = kept samples per second of whole epochs; `effective_samples_per_s`
= value / (1 + 2 tau) with the reference's binning tau
(aggregator.cpp:120-128).  Side lines in the same JSON object: the toric
4x4 workload in FP32 and FP64, the Pfaffian workload (Kitaev 4x4 + Pfaffian,
8192 chains), the same headline at N(0, 1e-4^2) weights, plain vs learned
proposal, `roofline` for the dominant kernel, `cpu_baseline` (the FP64 CPU
oracle on a bounded sample of the headline workload) and its calibration
against the reference's own per-rank numbers (toric n = 3, 4, 6 plogs).
"""
import argparse
import json
import os
import platform
import struct
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rbm-nqs_amd", "python"))

import numpy as np  # noqa: E402

METRIC = "samples/s and SR epoch ms at fixed energy error, plain vs learned proposal"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)
CHECKPOINT = os.path.join(ROOT, "rbm-nqs_amd", "examples", "kitaev6_sym.rbm")
E_EXACT_KITAEV6 = -0.78823570803603  # kitaev.cpp:66-100 at 6x6


def kitaev6_spec(chains=16384, spc=8, warmup=100, weights=0.05, learn=0, max_iter=100):
    s = f"""
[model]
type = kitaev
n_cells = 6
J = -1
symmetry = 0.5
[rbm]
type = symmetry
alpha = 4
weights = {weights}
[sampler]
n_samples = {chains * spc}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
metropolis.bond_flips = 0.5
[optimizer]
learning_rate = 0.01
sr.max_iterations = {max_iter}
"""
    if learn:
        s += f"[learn]\nhidden = {learn}\nfraction = 0.5\n"
    return s


def toric4_spec(chains=4096, spc=64, warmup=100):
    return f"""
[model]
type = toric
n_cells = 4
J = -1
[rbm]
type = basic
alpha = 2
weights = 0.05
[sampler]
n_samples = {chains * spc}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
metropolis.bond_flips = 0.5
[optimizer]
sr.max_iterations = 100
"""


def pfaffian_spec(chains=8192, spc=8, warmup=100):
    return f"""
[model]
type = kitaev
n_cells = 4
J = -1
symmetry = 0.5
[rbm]
type = basic
alpha = 2
weights = 0.05
pfaffian = true
pfaffian.weights = 0.01
[sampler]
n_samples = {chains * spc}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
metropolis.bond_flips = 0.5
pfaffian_refresh = 64
[optimizer]
sr.max_iterations = 100
"""


class Decay:
    """decay_t of abstract_optimizer.cpp:27-49 (decays before returning)."""

    def __init__(self, initial, minimum, decay):
        self.v, self.min, self.decay, self.done = initial, minimum, decay, decay == 1.0

    def get(self):
        if not self.done:
            self.v *= self.decay
            if self.v < self.min:
                self.v, self.done = self.min, True
        return self.v


def read_rbm(path, m):
    """a .rbm checkpoint (rbm_base.cpp:219-288, eigen_fstream.cpp:25-51) as
    the parameter vector [v | h | W column-major]"""
    data = open(path, "rb").read()
    off, mats = 0, []
    for _ in range(3):
        r, c = struct.unpack_from("<QQ", data, off)
        off += 16
        mats.append(np.frombuffer(data, dtype=np.complex128, count=r * c, offset=off).copy())
        off += 16 * r * c
    W, h, v = mats
    p = np.concatenate([v, h, W])
    assert p.size == m.P, (p.size, m.P)
    return p


def epochs(spec, precision, steps, warmup, params=None, graph=True, seed=421390484):
    """whole SR epochs on the device; returns timing and statistics"""
    import rbmgpu

    m = rbmgpu.Model(spec)
    g = m.gpu(precision=precision)
    g.set_params(m.initial_weights(seed) if params is None else params)
    lr, k1, k2, kd = Decay(1e-2, 1e-2, 1.0), Decay(1, 1e-4, 0.9), Decay(1e-3, 1e-6, 0.9), Decay(1e-2, 1e-2, 1.0)
    mi = 100
    log = []

    def one(e):
        st = g.train_epoch(e, lr.get(), k1.get(), k2.get(), kd.get(), max_iter=mi, graph=graph)
        log.append(st)

    for e in range(warmup):
        one(e)
    g.synchronize()
    t0 = time.perf_counter()
    for e in range(warmup, warmup + steps):
        one(e)
    g.synchronize()
    dt = time.perf_counter() - t0
    timed = log[warmup:]
    tau = float(np.mean([s["tau"] for s in timed]))
    S = m.n_samples
    out = {"ms_per_epoch": 1e3 * dt / steps, "samples_per_s": S * steps / dt, "samples_per_epoch": S,
           "tau_int": tau, "effective_samples_per_s": S * steps / dt / (1 + 2 * max(tau, 0.0)),
           "energy_per_site": timed[-1]["mean"].real / m.N, "energy_stddev_per_site": timed[-1]["std"] / m.N,
           "acceptance": float(np.mean([s["acc"] for s in timed])), "cg_iterations": timed[-1]["itn"],
           "graph_replayed": all(s["graph"] for s in timed) if graph else False,
           "cg_on_explicit_S": bool(timed[-1]["explicit_s"]),
           "launch_shape": list(g.launch_shape()), "structured_sx": g.structured()}
    return m, g, out


def sampler_flops(m, n_chains, sweeps, kept, bond_flips=0.5, n_offdiag_terms=None, flips_per_term=2):
    """algorithmic FP32 flops of one sampling epoch (the work model of the
    sampler): a slot proposes a bond (2 flips, probability bond_flips,
    consuming two slots) or a single flip, so a sweep of N slots makes
    N (1 - b) / (1 + b) single and N b / (1 + b) bond proposals; per unit
    24 flops for a single-flip ratio and 40 for a two-flip one (composition
    of the cosh / sinh factors, the ratio, the group product and the
    accepted-move update); each kept sample evaluates its off-diagonal
    local terms at 8 m flops per unit (m flipped sites)"""
    N, H = m.N, m.H
    b = bond_flips
    singles = N * (1 - b) / (1 + b)
    bonds = N * b / (1 + b)
    per_sweep = (24 * singles + 40 * bonds) * H
    nt = n_offdiag_terms if n_offdiag_terms is not None else N
    per_sample = nt * 8 * flips_per_term * H
    return n_chains * sweeps * per_sweep + n_chains * kept * per_sample


def sampler_roofline(g, m, chains, spc, warmup):
    """roofline of the dominant kernel of the headline epoch: the sampler
    (k_sample: all chains, warm-up and kept sweeps, in-sampler local
    energies).  It runs on the FP32 VALU: achieved = the algorithmic flops
    of sampler_flops / the kernel's duration between HIP events on the
    context's stream (an eager sampling epoch after the timed loop)."""
    ms = []
    for e in range(3):
        g.sample_epoch(epoch=900 + e)
        g.reduce(want_o=False)  # reads the epoch's event times
        ms.append(g.kernel_times()["sample"])
    t = float(np.median(ms))
    flops = sampler_flops(m, chains, warmup + spc, spc)
    achieved = flops / (t * 1e-3) / 1e12
    return {"kernel": "k_sample_wave (Metropolis chains of the epoch with their local energies)", "bound": "mfma",
            "issue": "FP32 VALU (157.3 TFLOP/s is MI355X's FP32 peak, vector and matrix alike, counting packed "
                     "FP32; without packed instructions the VALU peak is 78.6 TFLOP/s)",
            "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
            "frac_of_nonpacked_peak": achieved / FP32_NONPACKED_TFLOPS,
            "traffic": SAMPLER_TRAFFIC_BYTES, "flops_per_launch": flops, "launch_ms": t}


# HBM bytes per sampler launch of the headline workload from two rocprofv3
# --pmc runs of `bench.py --sampler-only 8` (FETCH_SIZE 1042.6 KiB x2.0,
# then WRITE_SIZE 299214.8 KiB x1.0, scaled like the k_struct figure below):
# the kept states and tanh rows it writes
SAMPLER_TRAFFIC_BYTES = int((1042.56 * 2.0 + 299214.81) * 1024)
# FP32 FMA peak without packed instructions: 256 CUs x 4 SIMDs x 16 lanes x 2 x 2.4 GHz
FP32_NONPACKED_TFLOPS = 78.6

# HBM bytes per k_struct launch of the toric 4x4 workload from two rocprofv3
# --pmc runs (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by
# the factor a plain 1 GiB device copy gives (FETCH x2.0, WRITE x1.0)
STRUCT_TRAFFIC_BYTES = int((61816.8 * 1048576 / 524302.4 + 3968.4) * 1024)


def struct_roofline(g, m, S):
    """the structured S.x of the toric 4x4 line (k_struct), against HBM:
    algorithmic bytes S (8 M + 8) + 16 P per launch"""
    sx_ms = g.time_sx(20)
    M = (m.P - m.N) // (m.N + 1)
    nbytes = S * (8 * M + 8) + 16 * m.P
    achieved = nbytes / (sx_ms * 1e-3) / 1e9 if sx_ms > 0 else 0.0
    return {"kernel": "k_struct", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": STRUCT_TRAFFIC_BYTES, "algorithmic_bytes": nbytes,
            "launch_ms": sx_ms}


ED2_CKPT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rbm-nqs_amd", "examples", "kitaev2_sym_ed.rbm")
ED2_INI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rbm-nqs_amd", "examples", "kitaev2_full.ini")


def ed2_spec(bond_flips, learn=0, lam=0.5, chains=256, spc=256, full=False):
    """the reference's Kitaev example (2x2, symmetric RBM alpha = 4) with a
    Metropolis sampler of chains x spc samples, or exact enumeration"""
    s = open(ED2_INI).read()
    if not full:
        s = s.replace("type = full", "type = metropolis")
    s = s.replace("n_samples = 1600", f"n_samples = {chains * spc}")
    s = s.replace("metropolis.n_chains = 4", f"metropolis.n_chains = {chains}")
    s = s.replace("metropolis.bond_flips = 0.2", f"metropolis.bond_flips = {bond_flips}")
    if learn:
        s += f"[learn]\nhidden = {learn}\nfraction = {lam}\n"
    return s


def learned_ab(epochs_=32, train_epochs=20, target=1e-5, chains=256, spc=256):
    """plain Metropolis against the learned cluster proposal on a state where
    plain moves mix badly: the reference's Kitaev 2x2 example trained to the
    exact-diagonalisation energy (rbm-nqs_amd/examples/kitaev2_sym_ed.rbm,
    E/N = -0.8660189 against ED -0.8660254), 256 chains x 256 samples per
    sampling epoch, 100 warm-up sweeps.  The error of each epoch's E/N is
    measured against the exact expectation value of the same weights (GPU
    exact enumeration, FP64); time to a fixed error = sampler time per epoch
    x (rms error / target)^2, the rms over 32 epochs (about 12 % relative
    uncertainty, 25 % on the time)."""
    import rbmgpu

    mf = rbmgpu.Model(ed2_spec(0.2, full=True))
    params = read_rbm(ED2_CKPT, mf)
    spec = lambda bf, **kw: ed2_spec(bf, chains=chains, spc=spc, **kw)  # noqa: E731
    gf = mf.gpu(precision=64)
    gf.set_params(params)
    gf.full_epoch()
    exact = gf.reduce(want_o=False)["mean"].real / mf.N
    del gf
    out = {"workload": "Kitaev 2x2 (N=8) symmetric RBM alpha=4 trained to the ED energy (the reference's "
                       f"examples/kitaev/kitaev.ini run), {chains} chains x {spc} samples, 100 warm-up sweeps, FP32",
           "exact_energy_per_site": exact, "target_error_per_site": target}

    def run(name, spec, train=0, lam_eval=None):
        m = rbmgpu.Model(spec)
        g = m.gpu(precision=32)
        g.set_params(params)
        tr, tms = None, []
        for e in range(train):
            g.sample_epoch(epoch=1000 + e)
            g.synchronize()
            t0 = time.perf_counter()
            tr = g.learn_train(lr=3e-3, beta=0.01, batch=m.n_samples)
            g.synchronize()
            tms.append((time.perf_counter() - t0) * 1e3)
        if lam_eval is not None:  # trained at the spec's fraction, measured at this one
            g.learn_set_fraction(lam_eval)
        ms, errs, accs, stds = [], [], [], []
        for e in range(epochs_):
            g.sample_epoch(epoch=e)
            r = g.reduce(want_o=False)
            ms.append(g.kernel_times()["sample"])
            errs.append(r["mean"].real / m.N - exact)
            accs.append(r["acc"])
            stds.append(r["std"] / m.N)
        t = float(np.mean(ms[1:]))
        rms = float(np.sqrt(np.mean(np.square(errs))))
        rec = {"sampler_ms_per_epoch": t, "rms_error_per_site": rms, "mean_error_per_site": float(np.mean(errs)),
               "binned_std_per_site": float(np.sqrt(np.mean(np.square(stds)))), "acceptance": float(np.mean(accs)),
               "ms_to_target_error": t * (rms / target) ** 2}
        if tr is not None:
            rec.update({"training_epochs": train, "learned_acceptance": tr["accepted"] / max(tr["proposed"], 1),
                        "train_ms_per_step": float(np.median(tms))})
        if lam_eval is not None:
            rec["learned_fraction"] = lam_eval
        out[name] = rec

    run("plain_bond_flips_0.2", spec(0.2))
    run("plain_bond_flips_0.5", spec(0.5))
    run("plain_single_flips", spec(0.0))
    run("learned_untrained", spec(0.2, learn=32, lam=0.5))
    run("learned", spec(0.2, learn=32, lam=0.5), train=train_epochs)
    # the north star's judge: plain single-flip Metropolis, and the learned
    # proposal with single flips as its plain fallback
    run("learned_single_fallback", spec(0.0, learn=32, lam=0.5), train=train_epochs)
    # the same network (trained at 0.5) measured at smaller learned fractions
    run("learned_single_fallback_lam0.2", spec(0.0, learn=32, lam=0.5), train=train_epochs, lam_eval=0.2)
    run("learned_single_fallback_lam0.1", spec(0.0, learn=32, lam=0.5), train=train_epochs, lam_eval=0.1)
    best_plain = min(out[k]["ms_to_target_error"] for k in out if k.startswith("plain"))
    out["learned_speedup_to_target_error_vs_best_plain"] = best_plain / out["learned"]["ms_to_target_error"]
    out["learned_speedup_to_target_error_vs_single_flips"] = (out["plain_single_flips"]["ms_to_target_error"] /
                                                               out["learned_single_fallback"]["ms_to_target_error"])
    return out


def plain_vs_learned(params, epochs_=3, train_epochs=5, target=1e-4):
    """headline model at the headline weights: plain Metropolis (single and
    bond flips) against the learned cluster proposal (72 -> 256 -> 72,
    bf16 MFMA, trained train_epochs REINFORCE/Adam steps on its own
    records first).  Time to a fixed error of E/N = epoch time x
    (std / target)^2."""
    import rbmgpu

    out = {"workload": "Kitaev 6x6 symmetric RBM alpha=4, 16384 chains x 8 samples, 100 warm-up sweeps, "
                       "headline weights, sampling epochs only", "target_error_per_site": target}
    for learn in (0, 256):
        m = rbmgpu.Model(kitaev6_spec(learn=learn))
        g = m.gpu(precision=32)
        g.set_params(params if params is not None else m.initial_weights(421390484))
        tr, tms = [], []
        if learn:
            for e in range(train_epochs):
                g.sample_epoch(epoch=1000 + e)
                g.synchronize()
                t0 = time.perf_counter()
                tr.append(g.learn_train(lr=3e-3, beta=0.01, batch=16384))
                g.synchronize()
                tms.append((time.perf_counter() - t0) * 1e3)
        ms, taus, stds, accs, es = [], [], [], [], []
        for e in range(epochs_):
            g.sample_epoch(epoch=e)
            r = g.reduce(want_o=False)
            ms.append(g.kernel_times()["sample"])
            taus.append(r["tau"])
            stds.append(r["std"] / m.N)
            accs.append(r["acc"])
            es.append(r["mean"].real / m.N)
        t = float(np.mean(ms))
        sd = float(np.sqrt(np.mean(np.square(stds))))
        rec = {"sampler_ms_per_epoch": t, "samples_per_s": m.n_samples / (t * 1e-3),
               "tau_int": float(np.mean(taus)), "acceptance": float(np.mean(accs)),
               "energy_per_site": float(np.mean(es)), "energy_stddev_per_site": sd,
               "ms_to_target_error": t * (sd / target) ** 2}
        if learn:
            rec.update({"hidden": learn, "learned_fraction": 0.5, "training_epochs": train_epochs,
                        "learned_acceptance": tr[-1]["accepted"] / max(tr[-1]["proposed"], 1),
                        "train_ms_per_step": float(np.median(tms))})
        out["learned" if learn else "plain"] = rec
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() else n


def cpu_baseline(params):
    """the FP64 CPU oracle (a port of the reference's algorithm: sequential
    chains per worker, per-proposal theta update + lncosh with the
    reference's state-keyed cache, bond flips, matrix-free CG) on the
    headline model and weights, bounded: 512 chains x 8 samples (the GPU's
    samples per chain and warm-up), chains and CG S.x on every host thread"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    import rbmgpu

    chains, spc = 512, 8
    spec = kitaev6_spec(chains, spc)
    o = po.Oracle(spec)
    o.set_params(params if params is not None else rbmgpu.Model(spec).initial_weights(421390484))
    threads = host_threads()
    t0 = time.perf_counter()
    st = o.epoch(seed=421390484, epoch=0, threads=threads, use_lut=True)
    dt = time.perf_counter() - t0
    S = chains * spc
    return dict(value=S / dt, unit="samples/s", cores=threads, kind="port", cpu_model=cpu_model(),
                effective_value=S / dt / (1 + 2 * max(st["tau"], 0.0)), tau_int=st["tau"],
                sample=f"1 SR epoch of the headline model and weights with {chains} chains x {spc} samples "
                       f"({S} samples, 100 warm-up sweeps, CG {int(st['itn'])} iterations); FP64 oracle with the "
                       f"reference's lncosh cache on {threads} threads (chains and CG); sampling "
                       f"{st['t_sample']:.2f} s, SR {st['t_opt']:.2f} s")


# the reference's own per-rank figures (examples/toric/backup/*.plog: symmetric
# RBM alpha = 1, model.symmetry = 1, one chain per rank, 5 sweeps per kept
# sample, about 200 warm-up sweeps per epoch): kept samples per second per
# rank of the Sampling region
PLOG = {3: (800, 11.23e3), 4: (800, 4.29e3), 6: (400, 1.24e3)}


def calibration():
    """the oracle on one thread with the plog settings, as a ratio to the
    reference's published per-rank sampling rate"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    out = {}
    for n, (samples, ref) in PLOG.items():
        spec = f"""
[model]
type = toric
n_cells = {n}
J = -1
symmetry = 1
[rbm]
type = symmetry
alpha = 1
weights = 0.01
[sampler]
n_samples = {samples}
metropolis.n_chains = 1
metropolis.n_steps_per_sample = 5
metropolis.n_warmup_steps = 200
metropolis.bond_flips = 0.2
[optimizer]
sr.max_iterations = 100
"""
        o = po.Oracle(spec)
        rng = np.random.default_rng(1)
        o.set_params(0.01 * (rng.normal(size=o.P) + 1j * rng.normal(size=o.P)))
        st = o.epoch(seed=1, epoch=0, threads=1, use_lut=True)
        rate = samples / st["t_sample"]
        out[f"toric_n{n}"] = {"oracle_samples_per_s_1core": rate, "reference_samples_per_s_per_rank": ref,
                              "ratio": rate / ref, "oracle_sr_ms": 1e3 * st["t_opt"]}
    return out


def copy_calibration(mib):
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def sampler_only(spc, workload):
    """sampling epochs alone (for counter runs under rocprofv3 --pmc)"""
    import rbmgpu

    spec = kitaev6_spec(spc=spc) if workload == "kitaev6" else toric4_spec(spc=spc)
    m = rbmgpu.Model(spec)
    g = m.gpu(precision=32)
    params = headline_params()[0] if workload == "kitaev6" else None
    g.set_params(m.initial_weights(421390484) if params is None else params)
    ms = []
    for e in range(3):
        t0 = time.perf_counter()
        g.sample_epoch(epoch=e)
        g.synchronize()
        ms.append(1e3 * (time.perf_counter() - t0))
    print(json.dumps({"sampler_only": {"workload": workload, "samples_per_chain": spc, "epoch_ms": ms,
                                       "launch_shape": list(g.launch_shape()), "wave_sampler": g.wave_sampler(),
                                       "acceptance": g.reduce(want_o=False)["acc"]}}))


def headline_params():
    import rbmgpu

    if os.path.exists(CHECKPOINT):
        return read_rbm(CHECKPOINT, rbmgpu.Model(kitaev6_spec())), "trained checkpoint kitaev6_sym.rbm"
    return None, "N(0, 0.05^2) re+im (rbm_base.cpp:68-88 initialiser, seed 421390484)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the plain vs learned proposal comparison")
    ap.add_argument("--no-side", action="store_true", help="headline line only")
    ap.add_argument("--eager", action="store_true", help="launch the epoch eagerly instead of replaying a graph")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--sampler-only", type=int, default=0, metavar="SPC")
    ap.add_argument("--workload", default="kitaev6", choices=["kitaev6", "toric4"])
    ap.add_argument("--learned-ab", action="store_true", help="only the plain vs learned comparison")
    a = ap.parse_args()
    if a.learned_ab:
        print(json.dumps({"plain_vs_learned": learned_ab()}))
        print(json.dumps({"plain_vs_learned_4096x64": learned_ab(chains=4096, spc=64)}))
        return
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.sampler_only:
        sampler_only(a.sampler_only, a.workload)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU, and the "
              "multi-GPU path is a design (chains shard across GPUs), not a run.", file=sys.stderr)
        sys.exit(2)
    params, wdesc = headline_params()
    m, g, main_line = epochs(kitaev6_spec(), 32, a.steps, a.warmup, params=params, graph=not a.eager)
    roof = sampler_roofline(g, m, 16384, 8, 100)
    out = {
        "metric": METRIC,
        "value": main_line["samples_per_s"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_line["ms_per_epoch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "Kitaev honeycomb 6x6 (N=72), symmetric RBM alpha=4 (72 symmetries, P=293), "
                        "16384 chains x 8 samples, 100 warm-up sweeps, bond_flips 0.5, FP32 sampler, SR+CG "
                        "(max 100 iterations, FP64 vectors; its mat-vec on the explicit S, built once per epoch "
                        "by an MFMA SYRK, when the cost model finds that cheaper); one step = one SR epoch "
                        "replayed from a HIP graph",
            "weights": wdesc,
            "samples_per_epoch": main_line["samples_per_epoch"],
            "parallelism": "single GPU",
        },
        "effective_samples_per_s": main_line["effective_samples_per_s"],
        "tau_int": main_line["tau_int"],
        "energy_per_site": main_line["energy_per_site"],
        "energy_stddev_per_site": main_line["energy_stddev_per_site"],
        "energy_error_vs_exact_per_site": main_line["energy_per_site"] - E_EXACT_KITAEV6,
        "acceptance": main_line["acceptance"],
        "cg_iterations": main_line["cg_iterations"],
        "cg_on_explicit_S": main_line["cg_on_explicit_S"],
        "graph_replayed": main_line["graph_replayed"],
        "roofline": roof,
    }
    if not a.no_side:
        side = {}
        _, _, side["kitaev6_weights_1e-4"] = epochs(kitaev6_spec(weights=1e-4), 32, a.steps, a.warmup)
        m2, g2, side["toric4_fp32"] = epochs(toric4_spec(), 32, a.steps, a.warmup)
        side["toric4_fp32"]["sx_roofline"] = struct_roofline(g2, m2, m2.n_samples)
        del g2
        _, _, side["toric4_fp64"] = epochs(toric4_spec(), 64, max(2, a.steps // 2), 1)
        _, _, side["kitaev4_pfaffian_fp32"] = epochs(pfaffian_spec(), 32, a.steps, a.warmup)
        _, _, side["kitaev6_eager"] = epochs(kitaev6_spec(), 32, a.steps, 1, params=params, graph=False)
        out["side"] = side
    if not a.no_learned:
        out["plain_vs_learned"] = learned_ab()
        out["c3_plain_vs_learned"] = plain_vs_learned(params)
    if not a.no_cpu:
        cb = cpu_baseline(params)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = {"samples_per_s": out["value"] / cb["value"],
                                 "effective_samples_per_s": out["effective_samples_per_s"] / cb["effective_value"]}
        out["cpu_calibration"] = calibration()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
