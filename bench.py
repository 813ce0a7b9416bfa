#!/usr/bin/env python3
"""Throughput of the phaseless-AFQMC walker step on one MI355X.

A step is one pass of the walker step over the whole population: force bias,
field draws (device counter-based RNG), VHS, Taylor propagation, Green's
function / overlap, phaseless weights, population control, QR every 5 steps,
and the local energy of every walker (exact, or surrogate + exact subset in the
control-variate path).  Inputs are resident in HBM before the timed region.

Workload (default): N=300 orbitals, 60+60 electrons, 1200 Cholesky vectors,
UHF trial, 4096 walkers, dt=0.005 -- synthetic Hamiltonian (vmc_amd.synthetic).

Prints one JSON line.  `python bench.py --gpus N` with N > 1 exits: this build
runs on a single GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = ("walker-steps/s/GPU mixed-prec+surrogate vs plain FP64, at ≤0.1 mHa energy shift")
PEAK_FP64_TFLOPS = 78.6


def exchange_flops(spec, nwalk):
    """Algorithmic FLOPs of one exchange launch: per walker and spin
    4 G n^2 N (f = R_g theta) + 8 G n^2 (sum_ij f_ij f_ji)."""
    N, G = spec["norb"], spec["nchol"]
    nsp = [spec["nalpha"]] if spec["walker_rhf"] else [spec["nalpha"], spec["nbeta"]]
    return nwalk * sum(4.0 * G * n * n * N + 8.0 * G * n * n for n in nsp)


def flops_per_walker_step(spec, ortho=5, energy=True):
    """Algorithmic real FLOPs of one walker-step (complex x real MAC = 4, complex x
    complex MAC = 8), E_L evaluated every step."""
    N, G = spec["norb"], spec["nchol"]
    T = N * (N + 1) / 2
    nsp = [spec["nalpha"]] if spec["walker_rhf"] else [spec["nalpha"], spec["nbeta"]]
    f = 0.0
    for n in nsp:
        f += 4 * n * n * N + 16 * n ** 3 + 8 * N * n * n      # overlap matrix, inverse, theta
        f += 4 * n * N * G                                     # force bias
        f += 48 * N * N * n                                    # 2 one-body + 5 Taylor products
        f += 4 * n * n * N + 6 * n ** 3                        # new overlap
        f += 16 * (2 * N * n * n - 2.0 / 3.0 * n ** 3) / ortho  # QR, amortised
    f += 4 * T * G                                             # VHS
    if energy:
        for n in nsp:
            f += 4 * G * n * n * N + 8 * G * n * n + 2 * G * n
            f += 4 * n * n * N + 16 * n ** 3 + 8 * N * n * n + 4 * N * n
    return f


STAGES = ["force_bias", "fields", "vhs", "onebody", "green", "weight", "ortho", "exchange", "comb", "taylor"]
PEAKS = {"fp64": 78.6, "bf16": 2500.0}


def run_path(ham, spec, total_steps, warmup, **kw):
    """Driver run with E_L every step (nsteps=1), device RNG; the first `warmup`
    steps are excluded from timing."""
    from vmc_amd.afqmc import AFQ_RNG_DEVICE, run_long_prop
    W = spec["nwalk"]
    w, e, res = run_long_prop(
        ham, ham.phi_up, ham.phi_dn, ham.phi_up, ham.phi_dn, seed=20261202, dt=spec["dt"], nsteps=1,
        nsweeps=total_steps + 1, ortho_steps=spec["ortho"], nwalk=W, nranks=1,
        walker_rhf=spec["walker_rhf"], cholesky_threshold=0.0, rng_mode=AFQ_RNG_DEVICE,
        warmup_steps=warmup, **kw)
    k = int(res.timed_steps)
    wall = float(res.timed_wall_s)
    ms = 1e3 * wall / k
    return dict(walker_steps_per_s=W * k / wall, ms_per_step=ms, steps=k, energies=np.array(e),
                weights=np.array(w), stage_ms_per_step={n: v / k for n, v in zip(STAGES, res.stage_ms)},
                est_var=float(res.est_var), err_per_sqrt_s=float(np.sqrt(res.est_var * ms * 1e-3)),
                train_rmse=float(res.train_rmse), valid_rmse=float(res.valid_rmse),
                cv_var_sub=float(res.cv_var_sub), cv_rms=float(res.cv_rms),
                cv_fallbacks=int(res.cv_fallbacks), train_samples=int(res.train_samples))


def stage_rooflines(spec, path, n_exact):
    """Algorithmic FLOP/s of each stage of one step (per launch for the kernels that
    launch several times) against the MFMA peak of the precision it runs in."""
    N, G, W = spec["norb"], spec["nchol"], spec["nwalk"]
    T = N * (N + 1) / 2
    nsp = [spec["nalpha"]] if spec["walker_rhf"] else [spec["nalpha"], spec["nbeta"]]
    st = path["stage_ms_per_step"]
    out = {}
    def add(name, kernel, flops, launches, prec):
        ms = st[name] / launches
        if ms <= 0:
            return
        ach = flops / launches / (ms * 1e-3) / 1e12
        out[name] = dict(kernel=kernel, launch_ms=ms, flops_per_launch=flops / launches,
                         achieved=ach, peak=PEAKS[prec], frac=ach / PEAKS[prec], unit="TFLOP/s",
                         bound="mfma", precision=prec)
    add("exchange", "k_exchange / k_exchange_bf16", exchange_flops(spec, n_exact), 1,
        "fp64" if path["exchange_prec"] == "fp64" else "bf16")
    add("taylor", "k_vmul (5 complex V t products, FP64 MFMA)", W * sum(40.0 * N * N * n for n in nsp), 5, "fp64")
    add("onebody", "k_expmul (E psi, FP64 MFMA)", W * sum(8.0 * N * N * n for n in nsp), 2 * len(nsp), "fp64")
    add("force_bias", "k_force_bias (FP64 MFMA)", W * sum(4.0 * n * N * G for n in nsp), 1, "fp64")
    add("vhs", "k_vhs (FP64 MFMA)", W * 4.0 * T * G, 1, "fp64")
    add("green", "k_green (FP64 MFMA + Gauss-Jordan)", W * sum(4 * n * n * N + 16 * n ** 3 + 8 * N * n * n
                                                              for n in nsp), 1, "fp64")
    return out


def cpu_baseline(ham, spec, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    orc = O.Oracle(ham, threads, spec["walker_rhf"])
    orc.prep_prop(ham.phi_up, ham.phi_dn, spec["dt"])
    d = ham.serial_det(ham.phi_up, ham.phi_dn)
    orc.set_walkers(np.tile(d, (threads, 1)))
    nsteps = max(1, int(os.environ.get("AFQ_CPU_STEPS", "8")))
    rate, secs = orc.bench(threads, nsteps, 1, spec["ortho"], 7)
    return dict(value=rate, unit="walker-steps/s", cores=threads, kind="port",
                sample="%d walkers x %d steps (same workload), E_L every step, QR every %d, "
                       "%.1f s on %d host threads (oracle/afqmc_oracle.cpp)"
                       % (threads, nsteps, spec["ortho"], secs, threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--train-steps", type=int, default=0,
                    help="exact FP64 steps whose (walker, E_K - E_J) pairs train the surrogate "
                         "(default: ~48k pairs)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--subset-inv", type=int, default=16)
    ap.add_argument("--risk-frac", type=float, default=0.05,
                    help="fraction of live walkers (largest |E_J - median|) always measured exactly")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the 1 GiB plain-copy kernel (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs one GPU holding one rank's walker population; "
              "--gpus %d is not supported" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    from vmc_amd.abi import AFQ_ENERGY_SURROGATE, AFQ_PREC_BF16X3
    from vmc_amd.synthetic import workload
    t0 = time.time()
    ham, spec = workload(a.workload)
    W = spec["nwalk"]
    print("[bench] workload %s built in %.1fs" % (a.workload, time.time() - t0), file=sys.stderr, flush=True)
    if a.calibrate:
        from vmc_amd.afqmc import debug_copy
        debug_copy(1 << 30)
    T = a.train_steps or max(2, (48 * 1024 + W - 1) // W)
    if a.fp64_only:
        T = 0
    total = T + a.warmup + a.steps
    fp64 = run_path(ham, spec, total, T + a.warmup)
    fp64["exchange_prec"] = "fp64"
    print("[bench] fp64: %.1f walker-steps/s, %.1f ms/step" % (fp64["walker_steps_per_s"], fp64["ms_per_step"]),
          file=sys.stderr, flush=True)
    paths = {"fp64": fp64}
    best = fp64
    if not a.fp64_only:
        mix = run_path(ham, spec, total, T + a.warmup, energy_mode=AFQ_ENERGY_SURROGATE,
                       prec_exchange=AFQ_PREC_BF16X3, surrogate_train_steps=T, surrogate_hidden=256,
                       surrogate_epochs=a.epochs, subset_inv=a.subset_inv, risk_frac=a.risk_frac)
        mix["exchange_prec"] = "bf16x3"
        print("[bench] mixed+surrogate: %.1f walker-steps/s, %.1f ms/step (train rmse %.3g, valid %.3g Ha)"
              % (mix["walker_steps_per_s"], mix["ms_per_step"], mix["train_rmse"], mix["valid_rmse"]),
              file=sys.stderr, flush=True)
        paths["mixed_surrogate"] = mix
        best = mix
    for name, pth in paths.items():
        nrisk = int(np.ceil(a.risk_frac * W))
        n_exact = W if name == "fp64" else nrisk + (W - nrisk + a.subset_inv - 1) // a.subset_inv
        pth["rooflines"] = stage_rooflines(spec, pth, n_exact)
    # the dominant kernel: largest share of the step
    shares = {k: best["stage_ms_per_step"][k] for k in best["rooflines"]}
    dom_name = max(shares, key=shares.get)
    dom = dict(best["rooflines"][dom_name])
    dom["stage"] = dom_name
    dom["traffic"] = None
    result_energy = {}
    if "mixed_surrogate" in paths:
        e64, emx = fp64["energies"], paths["mixed_surrogate"]["energies"]
        d = (emx - e64)[T + 1:]
        result_energy = dict(
            energy_shift_mha=1e3 * float(np.mean(d)), energy_shift_sem_mha=1e3 * float(np.std(d, ddof=1) / np.sqrt(len(d)))
            if len(d) > 1 else None, max_abs_step_shift_mha=1e3 * float(np.max(np.abs(d))), compared_steps=len(d),
            speedup_vs_fp64=paths["mixed_surrogate"]["walker_steps_per_s"] / fp64["walker_steps_per_s"],
            err_per_sqrt_s_ratio=paths["mixed_surrogate"]["err_per_sqrt_s"] / fp64["err_per_sqrt_s"]
            if fp64["err_per_sqrt_s"] > 0 else None)
        print("[bench] speedup %.2fx, energy shift %.4f mHa (sem %s), error/sqrt(s) ratio %s"
              % (result_energy["speedup_vs_fp64"], result_energy["energy_shift_mha"],
                 result_energy["energy_shift_sem_mha"], result_energy["err_per_sqrt_s_ratio"]),
              file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(ham, spec, threads)
        print("[bench] cpu baseline: %.2f walker-steps/s on %d threads" % (cpu["value"], threads),
              file=sys.stderr, flush=True)
    wstep = flops_per_walker_step(spec)
    for pth in paths.values():
        pth.pop("energies", None)
        pth.pop("weights", None)
        pth["step_tflops_algorithmic"] = wstep * pth["walker_steps_per_s"] / 1e12
    out = {
        "metric": METRIC,
        "value": best["walker_steps_per_s"],
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": best["steps"],
        "warmup": T + a.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+bf16x3" if "mixed_surrogate" in paths else "f64",
        "data": "synthetic",
        "config": {"workload": "N=%d %da+%db G=%d %s trial, %d walkers, dt=%g, E_L every step, QR every %d"
                   % (spec["norb"], spec["nalpha"], spec["nbeta"], spec["nchol"], spec["trial"].upper(),
                      W, spec["dt"], spec["ortho"]),
                   "name": a.workload, "path": "mixed_surrogate" if "mixed_surrogate" in paths else "fp64",
                   "subset_inv": a.subset_inv, "risk_frac": a.risk_frac, "surrogate_train_steps": T},
        "roofline": dom,
        "cpu_baseline": cpu,
        "paths": paths,
        "flops_per_walker_step": wstep,
    }
    out.update(result_energy)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
