#!/usr/bin/env python3
"""Benchmark: effective samples/sec including E_loc and O_k on one MI355X.

Workload (the plain-method GPU configuration): 1-D transverse-field Ising
chain N=20, periodic, J=h=1; RBM with M=80 hidden units, complex64 parameters
drawn i.i.d. N(0, 0.01^2) (seed 1236); 4096 Metropolis single-spin-flip
chains, 200 thermalisation sweeps and 1000 measurement sweeps (one sample per
sweep of N proposals).  One step is one ExpectationValue.gradient call: the
full sampling pass, E_loc of every sample, O_k = my_tanh(theta) and the
FP64 gradient contraction.

value = chains * samples_per_chain / (2 tau_int(E_loc)) / seconds_per_step,
with tau_int from Sokal's automatic windowing (c = 5) on the per-chain E_loc
series.  The roofline entry prices the dominant kernel (the Metropolis
kernel, FP32 vector ALU bound) with algorithmic FLOP from rbmongpu_amd/flops.py
over its average launch time measured with HIP events on the library stream.
The CPU baseline times the FP64 oracle (the reference's algorithm, with its
per-string O_k(s') estimator) on one host core over a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, M = 20, 80
CHAINS, SAMPLES_PER_CHAIN, THERM, SWEEPS = 4096, 1000, 200, 1
SEED = 1236
PEAK_FP32_TFLOPS = 157.3


def make_params():
    rng = np.random.default_rng(SEED)
    b = 0.01 * (rng.normal(size=M) + 1j * rng.normal(size=M))
    W = 0.01 * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    return b, W


def tau_int(series, c=5.0):
    """Sokal's automatic windowing on (T, chains) series: the autocovariance
    is averaged over chains, tau(W) = 1/2 + sum_{t=1}^W rho(t), and W is the
    smallest window with W >= c tau(W)."""
    x = series - series.mean()
    T = x.shape[0]
    f = np.fft.rfft(x, n=2 * T, axis=0)
    acov = np.fft.irfft(f * np.conj(f), axis=0)[:T].mean(axis=1) / T
    rho = acov / acov[0]
    tau = 0.5
    for w in range(1, T):
        tau += rho[w]
        if w >= c * tau:
            return float(tau)
    return float(tau)


def kernel_source_sha():
    """Hash of the sampler's source: a stored traffic measurement is used only
    for the kernel it was taken on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("rbm_kernels.hip", "rbm_device.hpp", "rbm_kernels.hpp"):
        h.update(open(os.path.join(ROOT, "rbmongpu_amd", "csrc", f), "rb").read())
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


def cpu_baseline(budget_s):
    """The oracle's Metropolis + E_loc + O_k gradient on one core, chains of
    the same workload (200 thermalisation sweeps, then measured sweeps)."""
    from oracle import oracle as O
    from rbmongpu_amd import hamiltonians as HAM
    b, W = make_params()
    psi = O.RBM(b, W)
    op = O.Operator(*HAM.tfim_1d(N))
    # calibrate: 1 chain, thermalisation + a few samples
    per = 4
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, per, SWEEPS, THERM, 1, seed=1), mode=0)
    t_small = time.perf_counter() - t0
    per2 = 24
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, per2, SWEEPS, THERM, 1, seed=2), mode=0)
    t_big = time.perf_counter() - t0
    t_sample = max((t_big - t_small) / (per2 - per), 1e-9)
    t_therm = max(t_small - per * t_sample, 0.0)
    # bounded run: as many whole chains (therm + samples) as fit the budget
    spc = max(1, int((budget_s / 4 - t_therm) / t_sample))
    chains = 4
    t0 = time.perf_counter()
    O.gradient(psi, op, O.MonteCarlo(N, spc * chains, SWEEPS, THERM, chains, seed=3), mode=0)
    wall = time.perf_counter() - t0
    return chains * spc / wall, chains, spc, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle CPU work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--samples-per-chain", type=int, default=SAMPLES_PER_CHAIN)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for the FETCH_SIZE/WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU: the sampler's chains are independent, and this build runs on a "
                         "single MI355X (multi-GPU sharding is described in DESIGN.md, not run)\n")
        sys.exit(2)

    import rbmongpu_amd as R
    from rbmongpu_amd import flops as F
    from rbmongpu_amd import hamiltonians as HAM

    spc = args.samples_per_chain
    b, W = make_params()
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=args.precision)
    op = R.Operator(*HAM.tfim_1d(N, 1.0, 1.0), True)
    mc = R.MonteCarloLoop(CHAINS * spc, SWEEPS, THERM, CHAINS, True)
    ev = R.ExpectationValue(True)
    for _ in range(args.warmup):
        ev.gradient(psi, op, mc)
    R.synchronize()
    R.timing_enable(True)
    R.timing_reset()
    R.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grad, E = ev.gradient(psi, op, mc)
    R.synchronize()
    wall = time.perf_counter() - t0
    R.timing_enable(False)
    ms_per_step = 1e3 * wall / args.steps
    kernel_ms = {k: R.timing_query(k) for k in ("mc", "eval", "coef", "scalar_reduce", "ok_reduce", "partials_sum")}
    mc_ms, mc_launches = kernel_ms["mc"]
    mc_avg_ms = mc_ms / max(mc_launches, 1)

    _, _, el = mc.last_samples()
    series = el.real.reshape(spc, CHAINS)
    tau = tau_int(series)
    samples = CHAINS * spc
    ess_per_s = samples / (2.0 * tau) / (ms_per_step / 1e3)

    terms = N  # N single-flip X terms; the N ZZ strings are diagonal
    launch_flops = F.mc_launch_flops(N, M, CHAINS, spc, THERM, SWEEPS)
    achieved = launch_flops / (mc_avg_ms / 1e3) / 1e12
    traffic = F.mc_launch_bytes(CHAINS, spc)
    ev_ms, ev_launches = kernel_ms["eval"]
    ev_avg_ms = ev_ms / max(ev_launches, 1)
    ev_flops = F.eval_flops(N, M, terms, CHAINS * spc)

    if args.copy_calibration:
        copy_calibration()

    traffic_meas = None
    tpath = os.path.join(ROOT, "profiles", "mc_rbm_c2_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("kernel_source_sha") == kernel_source_sha():
            traffic_meas = t["bytes_per_launch"]

    cpu = None
    if not args.no_cpu_baseline:
        cps, cchains, cspc, cwall = cpu_baseline(args.cpu_budget)
        cpu = {"value": cps / (2.0 * tau), "unit": "effective samples/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle (reference algorithm: per-string O_k(s') gradient estimator), {cchains} chains x "
                         f"({THERM} thermalisation + {cspc} measured sweeps) of the same N=20, M=80 TFIM workload in "
                         f"{cwall:.1f} s on one core; raw {cps:.1f} samples/s divided by 2 tau_int of the GPU run"}

    out = {
        "metric": "effective samples/sec incl. E_loc+O_k on 1 GPU (plain single-flip proposal)",
        "value": ess_per_s,
        "unit": "effective samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if args.precision == 32 else "c128",
        "data": "synthetic (seeded N(0, 0.01^2) RBM parameters)",
        "config": {"workload": "1-D TFIM N=20 J=h=1 periodic, RBM M=80, Metropolis single-spin-flip, "
                               f"{CHAINS} chains x ({THERM} therm + {spc} measurement sweeps), gradient step "
                               "(sampling + E_loc + O_k + FP64 gradient contraction)",
                   "chains": CHAINS, "samples_per_chain": spc, "thermalization_sweeps": THERM,
                   "N": N, "M": M, "precision": args.precision},
        "tau_int": tau,
        "samples_per_s": samples / (ms_per_step / 1e3),
        "acceptance_rate": mc.acceptance_rate,
        "energy": E.real,
        "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kernel_ms.items()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_meas,
                     "kernel": "k_mc_rbm", "flop_per_launch": launch_flops, "avg_launch_ms": mc_avg_ms,
                     "algorithmic_bytes_per_launch": traffic},
        "eval_kernel": {"kernel": "k_eval_rbm (log psi, E_loc, my_tanh of the recorded samples)",
                        "avg_launch_ms": ev_avg_ms, "flop_per_launch": ev_flops,
                        "achieved_tflops": ev_flops / (ev_avg_ms / 1e3) / 1e12 if ev_avg_ms > 0 else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
