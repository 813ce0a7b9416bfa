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


def run_path(ham, spec, steps, warmup, **prec):
    from vmc_amd.afqmc import AFQ_RNG_DEVICE, run_long_prop
    W = spec["nwalk"]
    w, e, res = run_long_prop(
        ham, ham.phi_up, ham.phi_dn, ham.phi_up, ham.phi_dn, seed=20261202, dt=spec["dt"], nsteps=1,
        nsweeps=warmup + steps + 1, ortho_steps=spec["ortho"], nwalk=W, nranks=1,
        walker_rhf=spec["walker_rhf"], cholesky_threshold=0.0, rng_mode=AFQ_RNG_DEVICE,
        warmup_steps=warmup, **prec)
    k = int(res.timed_steps)
    wall = float(res.timed_wall_s)
    return dict(walker_steps_per_s=W * k / wall, ms_per_step=1e3 * wall / k, steps=k,
                energies=e[warmup + 1:], weights=w[warmup + 1:], stage_ms=list(res.stage_ms),
                nchol_ene=int(res.nchol_ene))


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the 1 GiB plain-copy kernel (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs one GPU holding one rank's walker population; "
              "--gpus %d is not supported (multi-GPU sharding is described in DESIGN.md)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    from vmc_amd.synthetic import workload
    t0 = time.time()
    ham, spec = workload(a.workload)
    spec["nwalk"] = spec["nwalk"]
    print("[bench] workload %s built in %.1fs" % (a.workload, time.time() - t0), file=sys.stderr, flush=True)
    if a.calibrate:
        from vmc_amd.afqmc import debug_copy
        debug_copy(1 << 30)
    fp64 = run_path(ham, spec, a.steps, a.warmup)
    print("[bench] fp64: %.1f walker-steps/s, %.1f ms/step" % (fp64["walker_steps_per_s"], fp64["ms_per_step"]),
          file=sys.stderr, flush=True)
    best = fp64
    W = spec["nwalk"]
    xms = fp64["stage_ms"][7] / fp64["steps"]  # one exchange launch per step
    xflops = exchange_flops(spec, W)
    achieved = xflops / (xms * 1e-3) / 1e12
    roof = dict(bound="mfma", kernel="k_exchange (FP64 MFMA 16x16x4)", achieved=achieved,
                peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=achieved / PEAK_FP64_TFLOPS, traffic=None,
                launch_ms=xms, flops_per_launch=xflops)
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(ham, spec, threads)
        print("[bench] cpu baseline: %.2f walker-steps/s on %d threads" % (cpu["value"], threads),
              file=sys.stderr, flush=True)
    wstep = flops_per_walker_step(spec)
    out = {
        "metric": METRIC,
        "value": best["walker_steps_per_s"],
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": best["steps"],
        "warmup": a.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "N=%d %da+%db G=%d %s trial, %d walkers, dt=%g, E_L every step, QR every %d"
                   % (spec["norb"], spec["nalpha"], spec["nbeta"], spec["nchol"], spec["trial"].upper(),
                      W, spec["dt"], spec["ortho"]),
                   "name": a.workload, "norb": spec["norb"], "nalpha": spec["nalpha"],
                   "nbeta": spec["nbeta"], "nchol": spec["nchol"], "nwalk": W},
        "roofline": roof,
        "cpu_baseline": cpu,
        "paths": {"fp64": {"walker_steps_per_s": fp64["walker_steps_per_s"], "ms_per_step": fp64["ms_per_step"],
                           "stage_ms_per_step": dict(zip(
                               ["force_bias", "fields", "vhs", "propagate", "green", "weight", "ortho",
                                "exchange", "comb"], [x / fp64["steps"] for x in fp64["stage_ms"]])),
                           "step_tflops": wstep * fp64["walker_steps_per_s"] / 1e12}},
        "flops_per_walker_step": wstep,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
