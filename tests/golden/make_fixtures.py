#!/usr/bin/env python3
"""Regenerate the committed reference fixtures under tests/golden/.

Reads the reference's own AFQMC regression inputs and outputs
(tests/DQMC/{rhf_rhf,uhf_uhf,rhf_uhf,uhf_rhf,uhf_uhf_ui}) from the read-only
reference checkout and writes them as plain numpy arrays (no pickle):

  <case>.npz : header, hcore, hcore_mod, chol (or chol_up/chol_dn, ...),
               energy_core, rhf (N x N) and/or uhf (N x 2N) trial text,
               samples (nblocks x 2: weight, energy) from samples.ref,
               and the afqmc.json sampling keys as scalars.

HDF5 is read with the h5dump tool (binary little-endian dump of one dataset);
h5py is not available in this image.  The datasets are the ones the reference
reader consumes (utils/integral.cpp:589-661 for intType "r",
utils/integral.cpp:664-741 for intType "u"); text matrices follow readMat
(utils/input.cpp:376-383, row-major whitespace text).

Run in the build container only (it needs the reference checkout):
    python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/tests/DQMC"
H5DUMP = "/opt/conda/bin/h5dump"
CASES = ["rhf_rhf", "uhf_uhf", "rhf_uhf", "uhf_rhf", "uhf_uhf_ui"]
HERE = os.path.dirname(os.path.abspath(__file__))


def h5_dataset(path, name, dtype):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path],
                       check=True, stdout=subprocess.DEVNULL)
        return np.fromfile(out, dtype=dtype)


def h5_names(path):
    txt = subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True,
                         text=True).stdout
    return [ln.split('"')[1] for ln in txt.splitlines() if "DATASET" in ln]


def main():
    for case in CASES:
        d = os.path.join(REF, case)
        h5 = os.path.join(d, "FCIDUMP_chol")
        arrs = {}
        for name in h5_names(h5):
            dt = "<i8" if name == "header" else "<f8"
            arrs[name] = h5_dataset(h5, "/" + name, dt)
        nelec, norb, ms, nchol = (int(v) for v in arrs["header"])
        for txt, ncol in (("rhf.txt", norb), ("uhf.txt", 2 * norb)):
            p = os.path.join(d, txt)
            if os.path.exists(p):
                m = np.loadtxt(p).reshape(norb, ncol)
                arrs[txt.split(".")[0]] = m
        arrs["samples"] = np.loadtxt(os.path.join(d, "samples.ref")).reshape(-1, 2)
        cfg = json.load(open(os.path.join(d, "afqmc.json")))
        samp = cfg["sampling"]
        arrs["seed"] = np.int64(samp["seed"])
        arrs["dt"] = np.float64(samp["dt"])
        arrs["nsteps"] = np.int64(samp["nsteps"])
        arrs["nwalk"] = np.int64(samp["nwalk"])
        arrs["stochasticIter"] = np.int64(samp["stochasticIter"])
        arrs["choleskyThreshold"] = np.float64(samp["choleskyThreshold"])
        arrs["orthoSteps"] = np.int64(samp["orthoSteps"])
        arrs["left"] = np.array(cfg["wavefunction"]["left"])
        arrs["right"] = np.array(cfg["wavefunction"]["right"])
        arrs["intType"] = np.array(cfg["system"].get("intType", "r"))
        arrs["nranks"] = np.int64(4)  # tests/runDQMC.sh:6 runs mpirun -np 4
        out = os.path.join(HERE, case + ".npz")
        np.savez_compressed(out, **arrs)
        print(case, "N=%d nelec=%d ms=%d G=%d" % (norb, nelec, ms, nchol),
              "->", os.path.relpath(out, HERE), file=sys.stderr)


if __name__ == "__main__":
    main()
