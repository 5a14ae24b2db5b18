"""Benchmark: SR training epochs of the toric code 4x4 (N = 32 spins),
basic complex RBM alpha = 2 (P = 32 + 64 + 2048 = 2144), 4096 Metropolis
chains x 64 kept samples (262 144 samples per epoch), 100 warm-up sweeps,
bond_flips 0.5, matrix-free SR solved by CG with FP64 vectors
(sr.max_iterations = 100 as in examples/toric/toric.ini).

One step = one full epoch: sample (persistent chain kernel with fused local
energies) -> O assembly -> reduction -> F -> CG solve -> weight update, all
on one GPU.  Inputs (weights, tables) are resident in HBM before timing.

Prints one JSON line (see the driver contract) with `roofline` for the
dominant kernel (the S.x product of the CG) and `cpu_baseline` (the FP64
CPU oracle on a bounded sample of the same workload, rank 0 only).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rbm-nqs_amd", "python"))

import numpy as np  # noqa: E402

METRIC = "samples/s and SR epoch ms at fixed energy error, plain vs learned proposal"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(model="toric", n_cells=4, rbm="basic", alpha=2, chains=4096, samples_per_chain=64,
                warmup=100, bond_flips=0.5, sr_max_iterations=100, precision=32, weights_std=0.05)


def spec(chains, samples_per_chain, warmup=100, max_iter=100):
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
n_samples = {chains * samples_per_chain}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
metropolis.n_steps_per_sample = 1
metropolis.bond_flips = 0.5
[optimizer]
learning_rate = 0.01
sr.max_iterations = {max_iter}
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


def gpu_epochs(steps, warmup):
    import rbmgpu

    m = rbmgpu.Model(spec(WORKLOAD["chains"], WORKLOAD["samples_per_chain"]))
    g = m.gpu(precision=WORKLOAD["precision"])
    g.set_params(m.initial_weights(421390484))
    lr, k1, k2, kd = Decay(1e-2, 1e-2, 1.0), Decay(1, 1e-4, 0.9), Decay(1e-3, 1e-6, 0.9), Decay(1e-2, 1e-2, 1.0)
    dw = np.zeros(g.P, dtype=np.complex128)
    log = []

    def epoch(e):
        nonlocal dw
        g.sample_epoch(epoch=e)
        r = g.reduce()
        F = r["eo"] - r["mean"] * r["o_mean"].conj()
        rate = lr.get()
        r1, r2, rd = k1.get(), k2.get(), kd.get()
        dw, itn, rs = g.sr_solve(F, dw, r1, r2, rd, max_iter=WORKLOAD["sr_max_iterations"])
        dw *= rate
        g.update_weights(dw)
        log.append(dict(E=r["mean"].real / m.N, std=r["std"] / m.N, acc=r["acc"], tau=r["tau"], itn=itn))

    for e in range(warmup):
        epoch(e)
    g.synchronize()
    t0 = time.perf_counter()
    for e in range(warmup, warmup + steps):
        epoch(e)
    g.synchronize()
    dt = time.perf_counter() - t0
    times = g.kernel_times()
    return m, g, dt, log, times


def cpu_baseline():
    """FP64 CPU oracle epoch (same model, bounded: 64 chains x 32 samples)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    chains, spc = 512, 32
    o = po.Oracle(spec(chains, spc))
    rng = np.random.default_rng(421390484)
    o.set_params(WORKLOAD["weights_std"] * (rng.normal(size=o.P) + 1j * rng.normal(size=o.P)))
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    st = o.epoch(seed=421390484, epoch=0, threads=threads, use_lut=True)
    dt = time.perf_counter() - t0
    S = chains * spc
    return dict(value=S / dt, unit="samples/s", cores=threads, kind="port",
                sample=f"1 SR epoch of the same model with {chains} chains x {spc} samples "
                       f"({S} samples, 100 warm-up sweeps, CG 100 iterations); FP64 oracle with the "
                       f"reference's lncosh cache, chains on {threads} OpenMP threads, CG on 1 thread; "
                       f"sampling {st['t_sample']:.2f} s, SR {st['t_opt']:.2f} s")


# HBM bytes per k_struct launch of this workload from two rocprofv3 --pmc
# runs (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch, mean over the CG
# matvecs), each scaled by the factor a plain 1 GiB device copy gives
# (bench.py --copy-calibration 1024: FETCH_SIZE reads 524302 KiB for
# 1048576 KiB read -> x2.0; WRITE_SIZE reads 1048576 KiB -> x1.0):
# 61817 KiB x 2.0 + 3968 KiB x 1.0
STRUCT_TRAFFIC_BYTES = int((61816.8 * 1048576 / 524302.4 + 3968.4) * 1024)


def sx_roofline(g, m, S, sx_ms):
    """Roofline of one S.x of the CG.

    Structured (basic RBM, FP32): k_struct reads T (S x M complex FP32), one
    state word per sample and x once per launch -- the algorithmic bytes,
    S (8 M + 8) + 16 P.  Its two GEMMs (G = Sigma X, Z = Sigma^T diag(y)
    conj(T): 8 S N M flops, plus 26 S M + 8 S N for y and the column sums)
    run on bf16 MFMA with exact three-term splits (three MFMAs per product
    at 16x the FP32 MFMA rate), so HBM is the binding roofline;
    `fp32_equiv_tflops` is the FP32 flop rate of the same launch.
    Materialised O: one pass over O (P x S complex FP32), HBM bound."""
    M = (m.P - m.N) // (m.N + 1)  # P = N + M + N M for the basic RBM
    if g.structured():
        nbytes = S * (8 * M + 8) + 16 * m.P
        flops = 8 * S * m.N * M + 26 * S * M + 8 * S * m.N
        achieved = nbytes / (sx_ms * 1e-3) / 1e9 if sx_ms > 0 else 0.0
        return {"kernel": "k_struct (one S.x of the CG: G = Sigma X, Z = Sigma^T diag(y) conj(T))",
                "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": STRUCT_TRAFFIC_BYTES, "algorithmic_bytes": nbytes,
                "fp32_equiv_tflops": flops / (sx_ms * 1e-3) / 1e12 if sx_ms > 0 else 0.0,
                "flops_per_launch": flops, "launch_ms": sx_ms}
    o_bytes = g.P * S * 8
    achieved = o_bytes / (sx_ms * 1e-3) / 1e9 if sx_ms > 0 else 0.0
    return {"kernel": "k_sx_partial (one S.x of the CG over the materialised O)", "bound": "hbm",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "launch_ms": sx_ms}


def c3_spec(learn, chains=16384, spc=8, warmup=100, hidden=256, fraction=0.5):
    s = f"""
[model]
type = kitaev
n_cells = 6
J = -1
symmetry = 0.5
[rbm]
type = symmetry
alpha = 4
weights = 0.05
[sampler]
n_samples = {chains * spc}
metropolis.n_chains = {chains}
metropolis.n_warmup_steps = {warmup}
"""
    if learn:
        s += f"[learn]\nhidden = {hidden}\nfraction = {fraction}\n"
    return s


def plain_vs_learned(epochs=3, train_epochs=5, target=1e-4):
    """Kitaev 6x6 symmetric RBM alpha=4: plain Metropolis (single and bond
    flips) against the learned cluster proposal (72 -> 256 -> 72 network,
    bf16 MFMA, trained for train_epochs REINFORCE/Adam steps on its own
    records first).  Per epoch: sampler time, binned tau and std of E/N;
    time to a fixed error of E/N = epoch time x (std / target)^2."""
    import rbmgpu

    out = {"workload": "Kitaev honeycomb 6x6 (N=72), symmetric RBM alpha=4 (P=293), weights N(0,0.05^2), "
                       "16384 chains x 8 samples, 100 warm-up sweeps, FP32 sampler",
           "target_error_per_site": target}
    for learn in (False, True):
        m = rbmgpu.Model(c3_spec(learn))
        g = m.gpu(precision=32)
        g.set_params(m.initial_weights(421390484))
        tr = []
        if learn:
            for e in range(train_epochs):
                g.sample_epoch(epoch=1000 + e)
                tr.append(g.learn_train(lr=3e-3, beta=0.01, batch=16384))
        ms, taus, stds, accs, es = [], [], [], [], []
        for e in range(epochs):
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
            rec.update({"hidden": 256, "learned_fraction": 0.5, "training_epochs": train_epochs,
                        "learned_acceptance": tr[-1]["accepted"] / max(tr[-1]["proposed"], 1)})
        out["learned" if learn else "plain"] = rec
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


def sampler_only(spc):
    """Sampling epochs alone on the default workload's chains and model with
    spc kept samples per chain (for counter runs under rocprofv3 --pmc: the
    difference of two spc values isolates the kept-sample local energies)."""
    import rbmgpu

    m = rbmgpu.Model(spec(WORKLOAD["chains"], spc))
    g = m.gpu(precision=WORKLOAD["precision"])
    g.set_params(m.initial_weights(421390484))
    ms = []
    for e in range(3):
        t0 = time.perf_counter()
        g.sample_epoch(epoch=e)
        g.synchronize()
        ms.append(1e3 * (time.perf_counter() - t0))
    print(json.dumps({"sampler_only": {"samples_per_chain": spc, "chains": WORKLOAD["chains"],
                                       "epoch_ms": ms}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the plain vs learned proposal comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--sampler-only", type=int, default=0, metavar="SPC",
                    help="only run three sampling epochs with SPC kept samples per chain")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.sampler_only:
        sampler_only(a.sampler_only)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU, and the "
              "multi-GPU path is a design (chains shard across GPUs), not a run.", file=sys.stderr)
        sys.exit(2)
    m, g, dt, log, times = gpu_epochs(a.steps, a.warmup)
    S = WORKLOAD["chains"] * WORKLOAD["samples_per_chain"]
    ms = 1e3 * dt / a.steps
    value = S * a.steps / dt
    # dominant kernel alone: back-to-back launches on the last sample set,
    # between two HIP events on the context's stream
    sx_ms = g.time_sx(20)
    roofline = sx_roofline(g, m, S, sx_ms)
    times = dict(times)
    times.pop("unused", None)
    times["sx_main"] = sx_ms
    last = log[-1]
    tau = float(np.mean([l["tau"] for l in log[a.warmup:]]))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "toric code 4x4 (N=32), basic RBM alpha=2 (P=2144), weights N(0,0.05^2) re+im, "
                        "4096 chains x 64 samples, 100 warm-up sweeps, bond_flips 0.5, SR+CG (max 100 "
                        "iterations, FP64 vectors); one step = one SR epoch",
            "samples_per_epoch": S,
            "parallelism": "single GPU",
        },
        "effective_samples_per_s": value / (1 + 2 * max(tau, 0.0)),
        "tau_int": tau,
        "energy_per_site": last["E"],
        "energy_stddev_per_site": last["std"],
        "acceptance": last["acc"],
        "cg_iterations": last["itn"],
        "kernel_ms": times,
        "launch_shape": list(g.launch_shape()),
        "roofline": roofline,
    }
    if not a.no_learned:
        out["plain_vs_learned"] = plain_vs_learned()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
