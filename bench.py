#!/usr/bin/env python3
"""Benchmark: effective samples/sec including E_loc and O_k on one MI355X,
learned (MADE independent Metropolis-Hastings) against plain single-spin-flip
proposals.

Default workload (--config tfim10x10-learned, the learned-proposal configuration): 2-D
transverse-field Ising model on a 10x10 periodic lattice, J = 1, h = 3.04;
RBM with M = 200 hidden units, complex64 parameters drawn i.i.d. N(0, 0.01^2)
(seed 1238), then 50 Adam steps (lr 5e-3) on the energy gradient from 8192
chains x 10 samples.  Both samplers then run 8192 chains:
  plain    MonteCarloLoop: 200 thermalisation sweeps, one sample per sweep of N
           single-spin-flip proposals;
  learned  MADE 100 -> 256 -> 100 trained for 10 000 Adam steps (batch
           16384; lr 2e-3, then 4e-4, then 1e-4) on 4 x 10^6 samples of the
           plain chain, used as an independent Metropolis-Hastings proposal
           (20 thermalisation proposals, one sample per proposal);
           --made-training spec trains it as the workload states instead
           (2000 Adam steps at lr 1e-3, batch 4096, on 10^6 samples).
One step is one ExpectationValue.gradient call: sampling, E_loc of every
sample, O_k and the FP64 gradient contraction.  MADE training is timed
separately (made_train_s) and is not in the step.

value = chains * samples_per_chain / (2 tau_int(E_loc)) / seconds_per_step for
the learned sampler, with tau_int from Sokal's automatic windowing (c = 5) on
the per-chain E_loc series; the plain sampler is measured the same way in the
same run (plain.value), and the two energies are compared in units of their
combined standard error.

--config tfim20-plain measures the plain GPU configuration (1-D TFIM N = 20, M = 80,
4096 chains x (200 + 1000 sweeps)) instead; --config heisenberg6x6-sr one stochastic-
reconfiguration step of the 6x6 Heisenberg RBM (M = 144) with its errors
against an FP64 recomputation; --config j1j2-deep the deep-network tier (J1-J2 N = 40,
PsiDeep 40 -> 160 -> 80 on BF16 MFMA layers, 32768 chains) with the
BF16 / FP32 / FP64 forward comparison on 10^5 configurations.

roofline prices the dominant kernel of the measured step with algorithmic FLOP
from rbmongpu_amd/flops.py over its average launch time from HIP events on the
library stream.  cpu_baseline times the FP64 oracle (the reference's
algorithm, one host core) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
METRIC = "effective samples/sec incl. E_loc+O_k on 1 GPU: learned vs single-flip proposal"


def tau_int(series, c=5.0):
    """Sokal's automatic windowing on (T, chains) series: the autocovariance
    is averaged over chains, tau(W) = 1/2 + sum_{t=1}^W rho(t), and W is the
    smallest window with W >= c tau(W)."""
    x = series - series.mean()
    T = x.shape[0]
    f = np.fft.rfft(x, n=2 * T, axis=0)
    acov = np.fft.irfft(f * np.conj(f), axis=0)[:T].mean(axis=1) / T
    if acov[0] <= 0:
        return 0.5
    rho = acov / acov[0]
    tau = 0.5
    for w in range(1, T):
        tau += rho[w]
        if w >= c * tau:
            return float(tau)
    return float(tau)


def kernel_source_sha():
    """Hash of the kernels' sources with comments and blank space removed: a
    stored traffic measurement is used only for the code it was taken on."""
    import hashlib
    import re
    h = hashlib.sha256()
    for f in ("rbm_kernels.hip", "rbm_device.hpp", "rbm_kernels.hpp", "made_kernels.hip", "made_kernels.hpp",
              "deep_mfma.hip", "deep_kernels.hpp", "gram_kernels.hip", "linalg_kernels.hpp"):
        text = open(os.path.join(ROOT, "rbmongpu_amd", "csrc", f)).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        h.update(" ".join(text.split()).encode())
    return h.hexdigest()[:16]


def copy_calibration(nbytes=1 << 30):
    """A 1 GiB device-to-device hipMemcpy (the runtime's copy kernel reads and
    writes nbytes): the reference reading for the FETCH_SIZE / WRITE_SIZE passes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def stored_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed PMC summary, if it
    was taken on the current sources."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    if t.get("kernel_source_sha") != kernel_source_sha():
        return None
    return t.get("bytes_per_launch", {}).get(kernel_key)


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def rbm_params(N, M, seed):
    rng = np.random.default_rng(seed)
    b = 0.01 * (rng.normal(size=M) + 1j * rng.normal(size=M))
    W = 0.01 * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    return b, W


def chain_stats(el, samples_per_chain, chains):
    series = el.real.reshape(samples_per_chain, chains)
    per_chain = series.mean(axis=0)
    return tau_int(series), float(per_chain.mean()), float(per_chain.std(ddof=1) / np.sqrt(chains))


def time_steps(R, ev, psi, op, ens, steps, warmup, kernels):
    for _ in range(warmup):
        ev.gradient(psi, op, ens)
    R.synchronize()
    R.timing_enable(True)
    R.timing_reset()
    R.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g, E = ev.gradient(psi, op, ens)
    R.synchronize()
    wall = time.perf_counter() - t0
    R.timing_enable(False)
    kms = {k: R.timing_query(k) for k in kernels}
    return wall / steps, E, kms


def _cpu_chain(job):
    """One oracle chain in a worker process: the reference algorithm (Metropolis
    sweeps, E_loc, per-string O_k(s') gradient estimator) timed for spc
    samples, then the same chain's E_loc series (untimed) for its tau_int."""
    b, W, tab, N, therm, spc, seed = job
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    psi, op = O.RBM(b, W), O.Operator(*tab)
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, spc, 1, therm, 1, seed=seed), mode=0)
    wall = time.perf_counter() - t0
    _, _, el, _ = O.records(psi, op, O.MonteCarlo(N, spc, 1, therm, 1, seed=seed))
    return wall, el.real


def cpu_cores():
    """Host cores this process may use (at most 16, one GPU's share of the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_cores(psi_b, psi_W, tab, N, therm, budget_s, cores):
    """One independent oracle chain per host core (worker processes), each
    sized to about budget_s of work.  Returns (samples/s over all cores, cores,
    samples per chain, wall seconds, tau_int of E_loc from the chains' own
    series)."""
    import multiprocessing as mp
    from oracle import oracle as O
    psi, op = O.RBM(psi_b, psi_W), O.Operator(*tab)
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, 2, 1, therm, 1, seed=1), mode=0)
    t_small = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, 6, 1, therm, 1, seed=2), mode=0)
    t_sample = max((time.perf_counter() - t0 - t_small) / 4, 1e-9)
    t_therm = max(t_small - 2 * t_sample, 0.0)
    spc = max(8, int((budget_s - t_therm) / t_sample))
    jobs = [(psi_b, psi_W, tab, N, therm, spc, 100 + k) for k in range(cores)]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_chain, jobs)
    wall = time.perf_counter() - t0
    walls = [r[0] for r in res]
    series = np.stack([r[1] for r in res], axis=1)
    return cores * spc / max(walls), cores, spc, max(walls), tau_int(series), wall


def cpu_baseline(psi_b, psi_W, tab, N, therm, budget_s):
    """The oracle's Metropolis + E_loc + O_k gradient (the reference algorithm,
    FP64, per-string O_k(s') estimator) on one core: chains of the same
    workload (thermalisation, then measured sweeps), as many samples as fit
    the time budget."""
    from oracle import oracle as O
    psi = O.RBM(psi_b, psi_W)
    op = O.Operator(*tab)
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, 2, 1, therm, 1, seed=1), mode=0)
    t_small = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, 6, 1, therm, 1, seed=2), mode=0)
    t_big = time.perf_counter() - t0
    t_sample = max((t_big - t_small) / 4, 1e-9)
    t_therm = max(t_small - 2 * t_sample, 0.0)
    chains = 2
    spc = max(1, int((budget_s / chains - t_therm) / t_sample))
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, spc * chains, 1, therm, chains, seed=3), mode=0)
    wall = time.perf_counter() - t0
    return chains * spc / wall, chains, spc, wall


# ---------------------------------------------------------------------------
# proposal-network training: (samples per chain of training data, batch,
# phases of (Adam steps, learning rate)).  "tuned" (the default) is the schedule
# chosen by measurement; "spec" is the one the learned-proposal workload states
# (2000 Adam steps at lr 1e-3, batch 4096, on 10^6 plain-chain samples).
MADE_TRAINING = {
    "tuned": (492, 16384, ((4000, 2e-3), (4000, 4e-4), (2000, 1e-4))),
    "spec": (123, 4096, ((2000, 1e-3),)),
}


def run_tfim10x10_learned(args, R, HAM, F):
    L, N, M, C, H = 10, 100, 200, 8192, 256
    spc = args.samples_per_chain or 500
    tab = HAM.tfim_2d(L, L, 1.0, 3.04)
    b, W = rbm_params(N, M, 1238)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=args.precision)
    op = R.Operator(*tab, True)
    ev = R.ExpectationValue(True)
    # 50 Adam steps on the energy gradient (b and W; the first 2N entries are alpha, beta).
    # lr 5e-3 reaches E/N = -3.229 to -3.230 from every chain seed tried; at lr 1e-2 the
    # outcome depends on rounding-level differences of the gradient (one seed in three ends
    # at E/N = -2.5 with 33 % acceptance)
    log("learned-proposal workload: pre-training the RBM (50 Adam steps)")
    t0 = time.perf_counter()
    mc = R.MonteCarloLoop(C * 10, 1, 20, C, True, seed=11)
    p = psi.params
    m = np.zeros(p.shape, complex)
    v = np.zeros(p.shape)
    for t in range(1, 51):
        g, E_pre = ev.gradient(psi, op, mc)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * np.abs(g) ** 2
        p = p - 5e-3 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        p[:2 * N] = 0
        psi.params = p
    R.synchronize()
    pretrain_s = time.perf_counter() - t0

    kernels = ("mc", "eval", "coef", "scalar_reduce", "ok_reduce", "partials_sum", "made_sample", "proposal_eval", "imh")
    log("learned-proposal workload: plain single-flip steps")
    plain = R.MonteCarloLoop(C * spc, 1, 200, C, True, seed=21)
    p_ms, p_E, p_k = time_steps(R, ev, psi, op, plain, args.steps, args.warmup, kernels)
    p_spins, _, el = plain.last_samples()
    p_tau, p_mean, p_se = chain_stats(el, spc, C)
    p_tau_m = tau_int(abs_magnetisation(p_spins, N).reshape(spc, C))
    p_ess = C * spc / (2 * p_tau) / p_ms

    # training samples from the plain chain (4 x 10^6 tuned, 10^6 spec), then
    # the proposal network
    data_spc, batch, schedule = MADE_TRAINING[args.made_training]
    log("learned-proposal workload: training data and proposal network")
    R.synchronize()
    t0 = time.perf_counter()
    data_mc = R.MonteCarloLoop(C * data_spc, 1, 200, C, True, seed=31)
    ev(psi, op, data_mc)
    data, _, _ = data_mc.last_samples()
    data_s = time.perf_counter() - t0
    log(f"learned-proposal workload: {data.shape[0]} training samples")
    if args.made_params:
        # a network trained by an earlier run of this workload (profiling runs)
        made = R.MADE(N, H, params=np.load(args.made_params))
        made_s, losses = None, [float("nan"), float("nan")]
    else:
        made = R.MADE(N, H, seed=5)
        R.synchronize()
        log("learned-proposal workload: training the proposal network")
        t0 = time.perf_counter()
        # Adam's state carries over between the phases; the falling learning
        # rate takes the proposal's acceptance from 0.84 (2000 steps at 1e-3
        # on 10^6 samples) to 0.91
        losses = np.concatenate([made.train(data, steps, batch=batch, lr=lr, seed=6 + k)
                                 for k, (steps, lr) in enumerate(schedule)])
        R.synchronize()
        made_s = time.perf_counter() - t0
        if args.save_made_params:
            np.save(args.save_made_params, made.params)

    log("learned-proposal workload: learned-proposal steps")
    learned = R.LearnedMonteCarloLoop(made, C * spc, 20, C, True, seed=41)
    l_ms, l_E, l_k = time_steps(R, ev, psi, op, learned, args.steps, args.warmup, kernels)
    l_spins, _, el = learned.last_samples()
    l_tau, l_mean, l_se = chain_stats(el, spc, C)
    l_tau_m = tau_int(abs_magnetisation(l_spins, N).reshape(spc, C))
    l_ess = C * spc / (2 * l_tau) / l_ms
    # amortised: the proposal network's training data (plain-chain samples) and
    # its training are paid once and shared by every later learned step
    setup_s = data_s + (made_s or 0.0)
    per_step_ess = C * spc / (2 * l_tau)

    def amortised(steps):
        return steps * per_step_ess / (setup_s + steps * l_ms)
    breakeven = setup_s * p_ess / (l_ms * (l_ess - p_ess)) if l_ess > p_ess else None

    # dominant kernel of the learned step: E_loc / O_k evaluation of the samples
    ev_ms, ev_n = l_k["eval"]
    ev_avg = ev_ms / max(ev_n, 1)
    runs = learned.num_runs   # configurations measured per call: one per run of equal samples
    terms = N
    ev_flop = F.eval_flops(N, M, terms, runs)
    achieved = ev_flop / (ev_avg / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": l_ess,
        "unit": "effective samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": l_ms * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64" if args.precision == 32 else "c128",
        "data": "synthetic (seeded N(0, 0.01^2) RBM, 50 Adam steps at lr 5e-3; proposal network trained on the plain chain)",
        "config": {"workload": "2-D TFIM 10x10 periodic J=1 h=3.04, RBM M=200 after 50 Adam steps; learned "
                               "(MADE 100-256-100 independent MH) vs plain single-spin-flip, 8192 chains, "
                               f"{spc} samples per chain, gradient step (sampling + E_loc + O_k + FP64 contraction)",
                   "chains": C, "samples_per_chain": spc, "N": N, "M": M, "made_hidden": H,
                   "precision": args.precision},
        "speedup_vs_plain": l_ess / p_ess,
        "energy_agreement_sigma": abs(l_E.real - p_E.real) / float(np.hypot(l_se, p_se)),
        "learned": {"value": l_ess, "ms_per_step": l_ms * 1e3, "tau_int": l_tau, "tau_int_abs_m": l_tau_m,
                    "energy": l_E.real, "energy_se": l_se, "acceptance_rate": learned.acceptance_rate,
                    "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in l_k.items() if v[1]},
                    "made_train_s": made_s, "made_loss_first_last": [float(losses[0]), float(losses[-1])],
                    "made_training": {"schedule": args.made_training, "samples": int(data.shape[0]),
                                      "batch": batch,
                                      "phases": [{"steps": st, "lr": lr} for st, lr in schedule]}},
        "plain": {"value": p_ess, "ms_per_step": p_ms * 1e3, "tau_int": p_tau, "tau_int_abs_m": p_tau_m,
                  "energy": p_E.real, "energy_se": p_se, "acceptance_rate": plain.acceptance_rate,
                  "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in p_k.items() if v[1]}},
        "amortised": {"note": "learned ESS/s with the one-off cost of the proposal network (generating its "
                              f"{data.shape[0]} plain-chain training samples, {data_s:.2f} s, and training it, "
                              f"{made_s if made_s is not None else float('nan'):.2f} s) spread over the learned steps "
                              "that reuse it; breakeven_steps is where the learned chain's cumulative effective "
                              "samples overtake the plain chain's in the same wall time",
                      "setup_s": setup_s, "training_data_s": data_s, "made_train_s": made_s,
                      "value_over_steps": {str(k): amortised(k) for k in (args.steps, 50, 1000)},
                      "breakeven_steps": breakeven} if made_s is not None else None,
        "pretrain": {"seconds": pretrain_s, "energy_per_site": E_pre.real / N},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": stored_traffic("k_eval_rbm@tfim10x10-learned"),
                     "kernel": "k_eval_rbm (E_loc, log psi and my_tanh of one configuration per run)",
                     "flop_per_launch": ev_flop, "avg_launch_ms": ev_avg,
                     "algorithmic_bytes_per_launch": F.eval_bytes(M, runs)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_legs(psi, N, M, tab, 200, args.cpu_budget, p_tau,
                                       "plain single-spin-flip chain with the reference algorithm (per-string "
                                       "O_k(s') gradient estimator) on the same trained RBM and Hamiltonian")
    else:
        out["cpu_baseline"] = None
    return out


def popcount64(x):
    """Bits set in each uint64 (SWAR count)."""
    x = x - ((x >> np.uint64(1)) & np.uint64(0x5555555555555555))
    x = (x & np.uint64(0x3333333333333333)) + ((x >> np.uint64(2)) & np.uint64(0x3333333333333333))
    x = (x + (x >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    return ((x * np.uint64(0x0101010101010101)) >> np.uint64(56)).astype(np.int64)


def abs_magnetisation(spins, N):
    """|m| = |sum_i s_i| / N of (n, 2) spin words."""
    ups = np.zeros(len(spins), np.int64)
    for w, bits in ((0, min(N, 64)), (1, N - 64)):
        if bits <= 0:
            continue
        mask = np.uint64(0xFFFFFFFFFFFFFFFF if bits >= 64 else (1 << bits) - 1)
        ups += popcount64(spins[:, w] & mask)
    return np.abs(2 * ups - N) / N


def cpu_legs(psi, N, M, tab, therm, budget, gpu_tau, what):
    """The FP64 oracle on one core and with one chain per host core; the
    effective rate divides by 2 tau_int of the oracle chains' own E_loc series."""
    b, W = R_params_of(psi, N, M)
    log(f"CPU baseline: FP64 oracle, one core ({budget:.0f} s)")
    cps1, cchains, cspc, cwall = cpu_baseline(b, W, tab, N, therm, budget)
    cores = cpu_cores()
    log(f"CPU baseline: FP64 oracle, {cores} cores ({budget:.0f} s)")
    cpsn, _, spc_n, wall_n, tau_cpu, _ = cpu_baseline_cores(b, W, tab, N, therm, budget, cores)
    return {
        "value": cpsn / (2.0 * tau_cpu), "unit": "effective samples/s", "cores": cores, "kind": "port",
        "one_core": {"value": cps1 / (2.0 * tau_cpu), "cores": 1, "raw_samples_per_s": cps1},
        "tau_int": tau_cpu, "gpu_plain_tau_int": gpu_tau,
        "sample": f"FP64 oracle, {what}: {cores} worker processes with one chain each x ({therm} thermalisation + "
                  f"{spc_n} measured sweeps) in {wall_n:.1f} s ({cpsn:.2f} samples/s), and one core running "
                  f"{cchains} chains x {cspc} sweeps in {cwall:.1f} s ({cps1:.2f} samples/s); both divided by 2 "
                  f"tau_int of the oracle chains' own E_loc series"}


def R_params_of(psi, N, M):
    b = np.asarray(psi.b)
    W = np.asarray(psi.W).reshape(N, M)
    return b, W


def run_tfim20_plain(args, R, HAM, F):
    N, M, C, THERM = 20, 80, 4096, 200
    spc = args.samples_per_chain or 1000
    b, W = rbm_params(N, M, 1236)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=args.precision)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    op = R.Operator(*tab, True)
    mc = R.MonteCarloLoop(C * spc, 1, THERM, C, True)
    ev = R.ExpectationValue(True)
    kernels = ("mc", "eval", "coef", "scalar_reduce", "ok_reduce", "partials_sum")
    ms, E, kms = time_steps(R, ev, psi, op, mc, args.steps, args.warmup, kernels)
    _, _, el = mc.last_samples()
    tau, mean, se = chain_stats(el, spc, C)
    ess = C * spc / (2 * tau) / ms
    # At this chain count (one wavefront per SIMD) the library samples in two
    # launches per step and evaluates the first one's samples on a second
    # stream meanwhile; a launch's FLOP and bytes are its share of the step's.
    mc_ms, mc_n = kms["mc"]
    mc_avg = mc_ms / max(mc_n, 1)
    mc_lps = max(mc_n, 1) / max(args.steps, 1)
    flop = F.mc_launch_flops(N, M, C, spc, THERM, 1) / mc_lps
    achieved = flop / (mc_avg / 1e3) / 1e12
    ev_ms, ev_n = kms["eval"]
    ev_avg = ev_ms / max(ev_n, 1)
    ev_flops = F.eval_flops(N, M, N, C * spc) / (max(ev_n, 1) / max(args.steps, 1))
    out = {
        "metric": METRIC + " (plain single-flip configuration)",
        "value": ess, "unit": "effective samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64" if args.precision == 32 else "c128",
        "data": "synthetic (seeded N(0, 0.01^2) RBM parameters)",
        "config": {"workload": "1-D TFIM N=20 J=h=1 periodic, RBM M=80, Metropolis single-spin-flip, "
                               f"{C} chains x ({THERM} therm + {spc} measurement sweeps), gradient step "
                               "(sampling + E_loc + O_k + FP64 gradient contraction)",
                   "chains": C, "samples_per_chain": spc, "thermalization_sweeps": THERM, "N": N, "M": M,
                   "precision": args.precision},
        "tau_int": tau, "energy": E.real, "energy_se": se, "samples_per_s": C * spc / ms,
        "acceptance_rate": mc.acceptance_rate,
        "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kms.items()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": stored_traffic("k_mc_rbm@tfim20-plain"),
                     "kernel": "k_mc_rbm", "flop_per_launch": flop, "avg_launch_ms": mc_avg,
                     "launches_per_step": mc_lps, "algorithmic_bytes_per_launch": F.mc_launch_bytes(C, spc) / mc_lps},
        "eval_kernel": {"kernel": "k_eval_rbm (log psi, E_loc, my_tanh of the recorded samples; the first "
                                  "launch of a step runs beside the second sampler launch, so its time includes "
                                  "the shared CUs)",
                        "avg_launch_ms": ev_avg, "flop_per_launch": ev_flops,
                        "achieved_tflops": ev_flops / (ev_avg / 1e3) / 1e12 if ev_avg > 0 else None},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_legs(psi, N, M, tab, THERM, args.cpu_budget, tau,
                                       "the reference algorithm (per-string O_k(s') gradient estimator) on the "
                                       "same N=20, M=80 TFIM workload")
    else:
        out["cpu_baseline"] = None
    return out


def deep_c4_model(seed=1237):
    """J1-J2 chain N = 40 with a dense 40 -> 160 -> 80 PsiDeep: complex
    weights with real and imaginary parts N(0, 0.05^2 / fan_in), zero biases,
    final weights 1 (the summed output)."""
    from rbmongpu_amd.new_neural_network import deep_connections
    N, M, C = 40, [160, 80], [40, 160]
    rng = np.random.default_rng(seed)
    conns = [np.asarray(c, np.uint32) for c in deep_connections(N, M, C)]
    weights = [(0.05 / np.sqrt(c)) * (rng.normal(size=(c, m)) + 1j * rng.normal(size=(c, m))) for c, m in zip(C, M)]
    biases = [np.zeros(m, complex) for m in M]
    final = np.ones(M[-1], complex)
    return N, M, C, conns, biases, weights, final


def run_j1j2_deep(args, R, HAM, F):
    N, M, C, conns, biases, weights, final = deep_c4_model()
    prec = args.precision if args.precision in (16, 32) else 16
    CH, THERM = 32768, 50
    spc = args.samples_per_chain or 10

    def make(p):
        return R.PsiDeep(np.zeros(N), np.zeros(N), biases, conns, weights, final, 1.0, False, True, precision=p)

    psi = make(prec)
    tab = HAM.j1j2_chain(N, 1.0, 0.5)
    op = R.Operator(*tab, True)
    mc = R.MonteCarloLoop(CH * spc, 1, THERM, CH, True, seed=4)
    ev = R.ExpectationValue(True)
    mc_label = "deep_mfma_mc" if prec == 16 else "deep_mc"
    kernels = (mc_label, "deep_mfma_eval", "deep_grad", "deep_grad_rows", "deep_grad_gemm", "deep_ok", "coef", "ok_reduce",
               "partials_sum", "scalar_reduce")
    ms, E, kms = time_steps(R, ev, psi, op, mc, args.steps, args.warmup, kernels)
    _, _, el = mc.last_samples()
    tau, mean, se = chain_stats(el, spc, CH)
    ess = CH * spc / (2 * tau) / ms
    acc_init = mc.acceptance_rate
    # At the specified initialisation (weights N(0, 0.05^2 / fan_in)) log psi is
    # ~1e-5, so |psi|^2 is flat: every proposal is accepted and E_loc is
    # constant to 1e-7.  The same step on the same network with weights 12x
    # larger (N(0, 0.6^2 / fan_in): Re log psi spread by 0.7 over random
    # configurations, a single flip moving it by 0.4 on average) measures a
    # distribution with structure; it is reported beside the specified line,
    # with 50 samples per chain so that tau_int has a window to resolve.  (At
    # 20x the distribution is so peaked that 1.4 % of the proposals are
    # accepted.)
    SCALE_U, SPC_U = 12.0, 50
    log(f"deep-network workload: the same step with weights x{SCALE_U:g}")
    big = R.PsiDeep(np.zeros(N), np.zeros(N), biases, conns, [SCALE_U * w for w in weights], final, 1.0, False, True,
                    precision=prec)
    mc_u = R.MonteCarloLoop(CH * SPC_U, 1, THERM, CH, True, seed=24)
    ms_u, E_u, _ = time_steps(R, ev, big, op, mc_u, args.steps, args.warmup, kernels)
    _, _, el_u = mc_u.last_samples()
    tau_u, _, se_u = chain_stats(el_u, SPC_U, CH)
    unit_line = {"value": CH * SPC_U / (2 * tau_u) / ms_u, "ms_per_step": ms_u * 1e3, "tau_int": tau_u,
                 "energy": E_u.real, "energy_se": se_u, "acceptance_rate": mc_u.acceptance_rate,
                 "samples_per_chain": SPC_U, "weight_scale": SCALE_U,
                 "note": f"the same gradient step with every layer's weights scaled by {SCALE_U:g} and "
                         f"{SPC_U} samples per chain; value is its effective samples/s"}
    mc_ms, mc_n = kms[mc_label]
    mc_avg = mc_ms / max(mc_n, 1)
    flop = F.deep_mc_launch_flops(N, M, CH, (THERM + spc) * N)
    valu_flop = CH * ((THERM + spc) * N + 1) * F.deep_forward_valu_flops(M)
    # the sampler is bound by its FP32 vector work (complex my_logcosh of every
    # hidden unit per proposal; counters: VALU issue 82 %, the BF16 products a
    # small part of the matrix cores' time), so the roofline prices that work
    # against the FP32 vector peak; the matrix-core rate is reported beside it
    mfma_achieved = flop / (mc_avg / 1e3) / 1e12
    if prec == 16:
        achieved, peak = valu_flop / (mc_avg / 1e3) / 1e12, PEAK_FP32_TFLOPS
    else:
        achieved, peak = mfma_achieved, PEAK_FP32_TFLOPS
    # forward check: BF16 MFMA vs complex64 vs complex128 on 1e5 random configurations
    rng = np.random.default_rng(99)
    words = np.zeros((100000, 2), np.uint64)
    words[:, 0] = rng.integers(0, 2 ** N, size=100000, dtype=np.uint64)
    lps = {p: R.log_psi_batch(make(p), words) for p in (16, 32, 64)}

    def err(a, b):
        d = np.abs(lps[a] - lps[b])
        return {"max": float(d.max()), "median": float(np.median(d)), "rms": float(np.sqrt((d ** 2).mean()))}

    out = {
        "metric": METRIC + " (deep-network tier, plain single-flip chain)",
        "value": ess, "unit": "effective samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (MFMA inputs), fp32 accumulation" if prec == 16 else "c64",
        "data": "synthetic (seeded N(0, 0.05^2/fan_in) PsiDeep weights)",
        "config": {"workload": "1-D J1-J2 Heisenberg N=40 periodic, J1=1, J2=0.5; PsiDeep dense 40->160->80, "
                               f"log-cosh, summed output; {CH} chains x ({THERM} therm + {spc} measurement sweeps) "
                               "single-flip Metropolis with the full forward pass per proposal; gradient step "
                               "(sampling + E_loc + O_k + FP64 gradient contraction)",
                   "chains": CH, "samples_per_chain": spc, "thermalization_sweeps": THERM, "N": N,
                   "layers": M, "precision": prec},
        "tau_int": tau, "energy": E.real, "energy_se": se, "samples_per_s": CH * spc / ms,
        "acceptance_rate": acc_init,
        "structured_weights": unit_line,
        "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kms.items()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": stored_traffic("k_deep_mfma_mc@j1j2-deep") if prec == 16 else None, "kernel": "k_deep_mfma_mc"
                     if prec == 16 else "k_deep_mc", "flop_per_launch": valu_flop if prec == 16 else flop,
                     "avg_launch_ms": mc_avg,
                     "mfma_flop_per_launch": flop if prec == 16 else None,
                     "mfma_achieved": mfma_achieved if prec == 16 else None,
                     "mfma_frac_of_bf16_peak": mfma_achieved / PEAK_BF16_TFLOPS if prec == 16 else None,
                     "algorithmic_bytes_per_launch": F.mc_launch_bytes(CH, spc)},
        "forward_check": {"configurations": 100000, "bf16_vs_fp64": err(16, 64), "fp32_vs_fp64": err(32, 64),
                          "bf16_vs_fp32": err(16, 32), "max_abs_log_psi": float(np.abs(lps[64]).max())},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        model = O.Deep(N, biases, [np.asarray(c) for c in conns], weights, final)
        oop = O.Operator(*tab)
        t0 = time.perf_counter()
        O.gradient(model, oop, O.MonteCarlo(N, 1, 1, THERM, 1, seed=1), mode=0)
        t_one = time.perf_counter() - t0
        spc_cpu = max(1, int(args.cpu_budget / max(t_one, 1e-3)))
        t0 = time.perf_counter()
        O.gradient(model, oop, O.MonteCarlo(N, spc_cpu, 1, THERM, 1, seed=2), mode=0)
        wall = time.perf_counter() - t0
        cps = spc_cpu / wall
        out["cpu_baseline"] = {
            "value": cps / (2.0 * tau), "unit": "effective samples/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (reference algorithm: full forward per proposal and per string, per-string "
                      f"O_k(s') gradient estimator), 1 chain x ({THERM} thermalisation + {spc_cpu} measured sweeps) "
                      f"of the same model in {wall:.1f} s on one core; raw {cps:.2f} samples/s divided by "
                      f"2 tau_int of the GPU run"}
    else:
        out["cpu_baseline"] = None
    return out


def run_heisenberg6x6_sr(args, R, HAM, F):
    """6x6 Heisenberg, RBM M = 144: one stochastic-reconfiguration step
    (sampling, E_loc, O_k, gradient, centred S on FP32 MFMA with FP64
    accumulation, FP64 Cholesky solve of S + 1e-4 I).  Errors of S, the
    gradient and the update are stated against an FP64 recomputation from the
    same samples of a 2048-chain ensemble (8192 samples)."""
    N, M, CH, THERM, EPS = 36, 144, 16384, 100, 1e-4
    spc = args.samples_per_chain or 16
    b, W = rbm_params(N, M, 1235)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    tab = HAM.heisenberg_2d(6, 6)
    op = R.Operator(*tab, True)
    mc = R.MonteCarloLoop(CH * spc, 1, THERM, CH, True, seed=5)
    kernels = ("mc", "eval", "ok_reduce", "s_matrix", "s_gram", "cholesky", "cholesky_solve")
    for _ in range(args.warmup):
        R.sr_update(psi, op, mc, EPS)
    R.synchronize()
    R.timing_enable(True)
    R.timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        upd, E = R.sr_update(psi, op, mc, EPS)
    R.synchronize()
    ms = (time.perf_counter() - t0) / args.steps
    R.timing_enable(False)
    kms = {k: R.timing_query(k) for k in kernels}
    _, _, el = mc.last_samples()
    tau, mean, se = chain_stats(el, spc, CH)
    P = M + N * M
    n = CH * spc
    gram_ms, gram_n = kms["s_gram"]
    flop = F.s_matrix_flops(P, n)
    # the Gram product runs every FP32 product as six bfloat16 part products
    # on the matrix cores: priced as those products against the BF16 peak,
    # with the FP32-equivalent rate beside it
    fp32_equiv = flop / (gram_ms / args.steps / 1e3) / 1e12
    achieved = 6 * fp32_equiv
    # FP64 recomputation on the samples of one 2048-chain call with 4 samples
    # per chain: 8192 samples, more than the P parameters with an O_k, so
    # S + eps I is not rank-deficient
    from oracle import oracle as O
    chk = R.MonteCarloLoop(8192, 1, THERM, 2048, True, seed=6)
    chk_s, chk_g = R.MonteCarloLoop(chk), R.MonteCarloLoop(chk)
    upd_c, E_c = R.sr_update(psi, op, chk, EPS)
    S_gpu = R.get_S_matrix_centred(psi, chk_s)
    g_gpu, _ = R.ExpectationValue(True).gradient(psi, op, chk_g)
    spins, _, el_gpu = chk_g.last_samples()
    t_cpu = time.perf_counter()
    s = np.where((spins[:, :1] >> np.arange(N, dtype=np.uint64)) & np.uint64(1), 1.0, -1.0)
    T = O.tanh(b[None, :] + s @ W)
    Ok = np.concatenate([T, (s[:, :, None] * T[:, None, :]).reshape(len(s), N * M)], axis=1)
    El = O.local_energy(O.RBM(b, W), O.Operator(*tab), spins)
    Oc = Ok - Ok.mean(axis=0)
    S_ref = (Oc.conj().T @ Oc) / len(s)
    g_ref = (2.0 / len(s)) * ((El - El.mean())[:, None] * Ok.conj()).sum(axis=0)
    A = S_ref + EPS * np.eye(P)
    x_ref = np.linalg.solve(A, g_ref)
    t_cpu = time.perf_counter() - t_cpu
    # the same estimator with the kernel's E_loc separates the contraction's
    # error from E_loc's; the update's error against the first-order
    # propagation of the S and gradient errors shows the solve adds none
    g_with_gpu_eloc = (2.0 / len(s)) * ((el_gpu - el_gpu.mean())[:, None] * Ok.conj()).sum(axis=0)
    x = upd_c[2 * N:]
    predicted = np.linalg.solve(A, (g_gpu[2 * N:] - g_ref) - (S_gpu - S_ref) @ x_ref)

    def rel(a, b_):
        return float(np.max(np.abs(a - b_)) / np.max(np.abs(b_)))

    resid = np.linalg.norm((S_gpu + EPS * np.eye(P)) @ upd_c[2 * N:] - g_gpu[2 * N:]) / np.linalg.norm(g_gpu[2 * N:])
    out = {
        "metric": "stochastic-reconfiguration step (sampling, E_loc, O_k, gradient, S, Cholesky) on 1 GPU",
        "value": n / ms, "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 kernels, S at fp32 accuracy on bf16 MFMA (three-part split), fp64 accumulation and Cholesky",
        "data": "synthetic (seeded N(0, 0.01^2) RBM parameters)",
        "config": {"workload": f"2-D Heisenberg 6x6 periodic (72 bonds, 216 Pauli strings), RBM M=144, {CH} chains x "
                               f"({THERM} therm + {spc} measurement sweeps), SR step with S ({P}x{P} complex) on "
                               f"the BF16 matrix cores at FP32 accuracy (three-part split, six products) and the FP64 "
                               f"Cholesky solve of S + {EPS} I",
                   "chains": CH, "samples_per_chain": spc, "P": P, "eps": EPS},
        "tau_int": tau, "energy": E.real, "energy_se": se,
        "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kms.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS,
                     # HBM bytes of one Gram dispatch (one batch of 12 480 samples)
                     "traffic": stored_traffic("k_gram_lower_bf3@heisenberg6x6-sr"),
                     "kernel": "k_gram_lower_bf3 (S Gram, lower 128x128 tiles, FP32 products as six bfloat16 part "
                               "products)",
                     "flop_per_step": 6 * flop, "fp32_flop_per_step": flop, "fp32_equivalent_tflops": fp32_equiv,
                     "ms_per_step": gram_ms / args.steps},
        "fp64_recomputation": {
            "samples": len(s), "S_rel_err": rel(S_gpu, S_ref),
            "gradient_rel_err": rel(g_gpu[2 * N:], g_ref),
            "gradient_rel_err_contraction": rel(g_gpu[2 * N:], g_with_gpu_eloc),
            "local_energy_rel_err": rel(el_gpu, El),
            "update_rel_err": rel(x, x_ref),
            "update_vs_first_order_propagation": float(np.linalg.norm((x - x_ref) - predicted) / np.linalg.norm(predicted)),
            "cholesky_residual": float(resid),
            "note": "the gradient's error is the complex64 E_loc rounding amplified by the centring (sigma_E / <E> "
                    "is 1e-2 at this state); with the kernel's E_loc in the FP64 estimator the contraction alone "
                    "is gradient_rel_err_contraction.  The update's error is the first-order propagation of the S "
                    "and gradient errors through (S + eps I)^-1 to within update_vs_first_order_propagation"},
        "cpu_baseline": {
            "value": len(s) / t_cpu, "unit": "samples/s", "cores": cpu_cores(), "kind": "port",
            "sample": f"FP64 recomputation of the same SR step on {len(s)} of the samples: oracle E_loc (one core) "
                      "and numpy my_tanh / O_k rows, centred S, gradient and LAPACK solve of S + eps I (numpy's BLAS "
                      f"threads) in {t_cpu:.1f} s"} if not args.no_cpu_baseline else None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=("tfim10x10-learned", "tfim20-plain", "heisenberg6x6-sr", "j1j2-deep"),
                    default="tfim10x10-learned")
    ap.add_argument("--precision", type=int, default=None, choices=(16, 32, 64),
                    help="RBM workloads: 32 (default) or 64; j1j2-deep: 16 (BF16 MFMA, default) or 32")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle CPU work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--samples-per-chain", type=int, default=0)
    ap.add_argument("--made-params", default="", help="tfim10x10-learned: load the proposal network (.npy) instead of training it")
    ap.add_argument("--made-training", default="tuned", choices=tuple(MADE_TRAINING),
                    help="tfim10x10-learned: proposal-network training schedule (spec: 2000 Adam steps at lr 1e-3, "
                         "batch 4096, on 10^6 samples, as the workload states; tuned: the default)")
    ap.add_argument("--save-made-params", default="", help="tfim10x10-learned: save the trained proposal network (.npy)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for the FETCH_SIZE/WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU: the chains are independent, and the library runs on a single "
                         "MI355X (sharding across GPUs is a design note, not run)\n")
        sys.exit(2)

    import rbmongpu_amd as R
    from rbmongpu_amd import flops as F
    from rbmongpu_amd import hamiltonians as HAM

    if args.precision is None:
        args.precision = 16 if args.config == "j1j2-deep" else 32
    run = {"tfim10x10-learned": run_tfim10x10_learned, "tfim20-plain": run_tfim20_plain,
           "heisenberg6x6-sr": run_heisenberg6x6_sr, "j1j2-deep": run_j1j2_deep}[args.config]
    out = run(args, R, HAM, F)
    if args.copy_calibration:
        copy_calibration()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
