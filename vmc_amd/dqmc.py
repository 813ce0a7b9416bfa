"""The reference DQMC binary's phaseless run on its own input directory.

Reads what executables/DQMC.cpp reads -- the JSON input (utils/input.cpp:61-276,
the keys the phaseless walk uses, with the reference's defaults; keys are
case-insensitive as in its boost iptree), the FCIDUMP_chol integrals
(utils/integral.cpp:589-661, or :664-741 for system.intType "u") and the rhf.txt / uhf.txt trial text
(utils/input.cpp:376-383), and spinRDM.0.0.txt where present (the background
density, DQMC/DQMCWalker.cpp:42-58) -- picks the walker type and trials as
executables/DQMC.cpp:74-185 does, runs calcMixedEstimatorLongProp on the GPU
(afq_run_long_prop) and writes samples.dat (DQMC/MixedEstimator.cpp:678-683)
and, with print.writeOneRDM, the 1-RDM files (:689-697).

    python -m vmc_amd.dqmc <dir> [afqmc.json] [--ranks 4] [--fp64]
"""
import argparse
import json
import os
import time
from dataclasses import dataclass

import numpy as np

from .abi import HamArrays
from .formats import (read_fcidump_chol, read_spin_rdm, read_text_matrix, trial_orbitals, walker_is_rhf,
                      write_one_rdm, write_samples)


@dataclass
class Settings:
    integrals: str = "FCIDUMP"
    int_type: str = "r"
    left: str = "rhf"
    right: str = "rhf"
    seed: int = 0
    phaseless: bool = False
    dt: float = 0.1
    nsteps: int = 10
    nwalk: int = 100
    stochastic_iter: int = 10000
    burn_iter: int = 0
    ortho_steps: int = 50
    ene0_guess: float = 1e10
    cholesky_threshold: float = 0.005
    weight_cap: float = -1.0
    write_one_rdm: bool = False
    scratch_dir: str = "./"


def _lower_keys(d):
    return {str(k).lower(): (_lower_keys(v) if isinstance(v, dict) else v) for k, v in d.items()}


def read_afqmc_json(path):
    """The keys of utils/input.cpp:61-276 that the phaseless walk reads, with the
    reference's defaults (sampling.seed defaults to the clock, as getTime())."""
    cfg = _lower_keys(json.load(open(path)))

    def get(key, default):
        node = cfg
        for part in key.lower().split("."):
            if not isinstance(node, dict) or part not in node:
                return default
            node = node[part]
        return node

    s = Settings()
    s.integrals = get("system.integrals", s.integrals)
    s.int_type = get("system.intType", s.int_type)
    s.left = str(get("wavefunction.left", s.left)).lower()
    s.right = str(get("wavefunction.right", s.right)).lower()
    s.seed = int(get("sampling.seed", int(time.time())))
    s.phaseless = bool(get("sampling.phaseless", s.phaseless))
    s.dt = float(get("sampling.dt", s.dt))
    s.nsteps = int(get("sampling.nsteps", s.nsteps))
    s.nwalk = int(get("sampling.nwalk", s.nwalk))
    s.stochastic_iter = int(get("sampling.stochasticIter", s.stochastic_iter))
    s.burn_iter = int(get("sampling.burnIter", s.burn_iter))
    s.ortho_steps = int(get("sampling.orthoSteps", s.ortho_steps))
    s.ene0_guess = float(get("sampling.ene0Guess", s.ene0_guess))
    s.cholesky_threshold = float(get("sampling.choleskyThreshold", s.cholesky_threshold))
    s.weight_cap = float(get("sampling.weightCap", s.weight_cap))
    s.write_one_rdm = bool(get("print.writeOneRDM", s.write_one_rdm))
    s.scratch_dir = str(get("print.scratchDir", s.scratch_dir))
    return s


def run_directory(directory, input_file="afqmc.json", *, nranks=1, ref_compat=True, max_steps=0,
                  write=True, **run_kw):
    """One phaseless run of the reference's input directory on the GPU.  nranks
    virtual ranks stand for `mpirun -np nranks` (nwalk walkers and seed + r each);
    ref_compat keeps the reference's FP32 Cholesky in the VHS (the arithmetic its
    samples.ref was made with).  Returns (weights, energies, RunResult, 1-RDM or None)."""
    from .afqmc import run_long_prop

    s = read_afqmc_json(os.path.join(directory, input_file))
    if s.int_type not in ("r", "u"):
        raise NotImplementedError("intType %r: restricted or unrestricted integrals only" % s.int_type)
    if not s.phaseless:
        raise NotImplementedError("only the phaseless walk (sampling.phaseless) is on this path")
    if s.left not in ("rhf", "uhf") or s.right not in ("rhf", "uhf"):
        raise NotImplementedError("trials %s/%s: only rhf and uhf are on this path" % (s.left, s.right))
    d = read_fcidump_chol(os.path.join(directory, s.integrals), s.int_type)
    N, na, nb = d.norb, d.nalpha, d.nbeta
    mats = {}
    for kind, ncol in (("rhf", N), ("uhf", 2 * N)):
        p = os.path.join(directory, kind + ".txt")
        if os.path.exists(p):
            mats[kind] = read_text_matrix(p, N, ncol)
    wrhf = walker_is_rhf(na, nb, s.right)
    lu, ld = trial_orbitals(s.left, mats, N, na, nb)
    ham = HamArrays(N, na, nb, d.ecore, d.h1, d.h1_mod, d.chol, lu, ld, trial_rhf=(s.left == "rhf"),
                    h1_dn=d.h1_dn, h1_mod_dn=d.h1_mod_dn, chol_dn=d.chol_dn)
    # reference dets: rhf.txt for an RHF walker, uhf.txt otherwise (DQMC/MixedEstimator.cpp:338-351)
    ru, rd = trial_orbitals("rhf" if wrhf else "uhf", mats, N, na, nb)
    iu, idn = trial_orbitals(s.right, mats, N, na, nb)
    # prepPropR takes its background density from spinRDM.0.0.txt when the file exists
    # (DQMC/DQMCWalker.cpp:42-58)
    bg = None
    p = os.path.join(directory, "spinRDM.0.0.txt")
    if os.path.exists(p):
        bg = np.array(read_spin_rdm(p, N, na + nb))
    rdm = None
    if s.write_one_rdm:
        rdm = np.zeros((N, N) if s.left == "rhf" else (2, N, N))
    w, e, res = run_long_prop(ham, ru, rd, iu, idn, seed=s.seed, dt=s.dt, nsteps=s.nsteps, nsweeps=s.stochastic_iter,
                              ortho_steps=s.ortho_steps, nwalk=s.nwalk, nranks=nranks, walker_rhf=wrhf,
                              ref_compat=ref_compat, cholesky_threshold=s.cholesky_threshold,
                              burn_iter=s.burn_iter, ene0_guess=s.ene0_guess,
                              weight_cap=s.weight_cap if s.weight_cap > 0 else 0.0, max_steps=max_steps,
                              one_rdm=rdm, background_rdm=bg, **run_kw)
    if write:
        write_samples(os.path.join(directory, "samples.dat"), w, e)
        if rdm is not None:
            write_one_rdm(os.path.join(directory, s.scratch_dir), rdm, res.one_rdm_weight)
    return w, e, res, rdm


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("directory")
    ap.add_argument("input", nargs="?", default="afqmc.json")
    ap.add_argument("--ranks", type=int, default=1, help="virtual MPI ranks (mpirun -np)")
    ap.add_argument("--fp64", action="store_true", help="FP64 Cholesky in the VHS instead of the reference's FP32")
    a = ap.parse_args()
    w, e, res, _ = run_directory(a.directory, a.input, nranks=a.ranks, ref_compat=not a.fp64)
    print("blocks %d, last energy %.10f, reference energy %.10f" % (len(e), e[-1], res.ref_energy))


if __name__ == "__main__":
    main()
