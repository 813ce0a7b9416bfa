"""Reference output formats (CPU): the 1-RDM files match what std::ostream writes
for the same numbers (writeOneRDM, DQMC/MixedEstimator.cpp:52-112: default
stream format, entries followed by two spaces), and samples.dat matches
boost::format("%.7e      %.10e \\n") (DQMC/MixedEstimator.cpp:681), which is C's printf."""
import json
import subprocess

import numpy as np

from vmc_amd.formats import write_one_rdm, write_samples

CXX_SRC = r"""
#include <fstream>
#include <iostream>
int main(int argc, char** argv) {
  std::ofstream f(argv[1]);
  double cw; int n;
  std::cin >> cw >> n;
  f << cw << std::endl;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) { double v; std::cin >> v; f << v << "  "; }
    f << std::endl;
  }
}
"""


def test_one_rdm_file_matches_ostream(tmp_path):
    rng = np.random.default_rng(3)
    N = 5
    rdm = rng.standard_normal((2, N, N)) * np.array([1e-7, 1.0, 123.456, 1e6, 3.0])[None, None, :]
    rdm[0, 1, 1] = 0.0
    cw = 7204.0273586439358
    exe = tmp_path / "w"
    (tmp_path / "w.cpp").write_text(CXX_SRC)
    subprocess.run(["g++", "-O1", str(tmp_path / "w.cpp"), "-o", str(exe)], check=True)
    paths = write_one_rdm(str(tmp_path), rdm, cw, rank=3)
    assert [p.split("/")[-1] for p in paths] == ["rdm_up_3.dat", "rdm_dn_3.dat"]
    for p, m in zip(paths, rdm):
        ref = tmp_path / "ref.dat"
        inp = "%r %d\n" % (cw, N) + " ".join(repr(float(v)) for v in m.ravel())
        subprocess.run([str(exe), str(ref)], input=inp, text=True, check=True)
        assert open(p).read() == ref.read_text()
    (single,) = write_one_rdm(str(tmp_path), rdm[0] + rdm[1], cw)
    assert single.endswith("/rdm_0.dat")


def test_samples_file_format(tmp_path):
    p = tmp_path / "samples.dat"
    write_samples(str(p), [80.0, 79.51234567], [-75.9839842460, -76.1])
    assert p.read_text() == ("8.0000000e+01      -7.5983984246e+01 \n"
                             "7.9512346e+01      -7.6100000000e+01 \n")


# ---------------------------------------------------------------- FCIDUMP_chol (HDF5)
import os  # noqa: E402
import shutil  # noqa: E402

import pytest  # noqa: E402

from vmc_amd.formats import read_fcidump_chol, write_fcidump_chol  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_TESTS = "/root/reference/tests/DQMC"   # read only when present (this container), never on a GPU run
io_built = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(GOLD), "..", "vmc_amd", "libvmc_fcidump.so")),
    reason="libvmc_fcidump.so not built (no HDF5 C library)")


@io_built
def test_fcidump_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    N, G = 7, 5
    h1 = rng.standard_normal((N, N))
    h1m = rng.standard_normal((N, N))
    L = rng.standard_normal((G, N, N))
    L = L + L.transpose(0, 2, 1)
    p = str(tmp_path / "FCIDUMP_chol")
    write_fcidump_chol(p, h1, h1m, L, nelec=9, ecore=3.25, ms=1)
    d = read_fcidump_chol(p)
    assert (d.nelec, d.norb, d.ms, d.nchol, d.nalpha, d.nbeta) == (9, N, 1, G, 5, 4)
    assert np.array_equal(d.h1, h1) and np.array_equal(d.h1_mod, h1m) and np.array_equal(d.chol, L)
    assert d.ecore == 3.25
    if shutil.which("h5dump") or os.path.exists("/opt/conda/bin/h5dump"):
        exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
        out = subprocess.run([exe, "-H", p], capture_output=True, text=True, check=True).stdout
        # the reference writer's types and shapes (h5py from numpy int64 / float64 arrays)
        assert 'DATASET "header"' in out and "H5T_STD_I64LE" in out and "( 4 ) / ( 4 )" in out
        assert 'DATASET "chol"' in out and "( %d ) / ( %d )" % (G * N * N, G * N * N) in out
        assert "SCALAR" in out


@io_built
def test_fcidump_unrestricted_round_trip(tmp_path):
    from vmc_amd.formats import write_fcidump_chol_u
    rng = np.random.default_rng(5)
    N, G = 4, 3
    h1 = [rng.standard_normal((N, N)) for _ in range(2)]
    h1m = [rng.standard_normal((N, N)) for _ in range(2)]
    L = [rng.standard_normal((G, N, N)) for _ in range(2)]
    p = str(tmp_path / "FCIDUMP_chol")
    write_fcidump_chol_u(p, h1, h1m, L, nelec=5, ecore=-1.5, ms=1)
    d = read_fcidump_chol(p, "u")
    assert (d.nelec, d.norb, d.nchol, d.nalpha, d.nbeta) == (5, N, G, 3, 2)
    assert np.array_equal(d.h1, h1[0]) and np.array_equal(d.h1_dn, h1[1])
    assert np.array_equal(d.h1_mod, h1m[0]) and np.array_equal(d.h1_mod_dn, h1m[1])
    assert np.array_equal(d.chol, L[0]) and np.array_equal(d.chol_dn, L[1]) and d.ecore == -1.5
    with pytest.raises(RuntimeError):
        read_fcidump_chol(p)   # no restricted datasets in an unrestricted file


@io_built
def test_fcidump_missing_file(tmp_path):
    with pytest.raises(RuntimeError):
        read_fcidump_chol(str(tmp_path / "absent"))


@io_built
@pytest.mark.parametrize("name", ["rhf_rhf", "uhf_uhf", "rhf_uhf", "uhf_rhf", "uhf_uhf_ui"])
def test_fcidump_reads_reference_files(name):
    """The reference's own FCIDUMP_chol files, read directly, equal the committed fixtures
    (which h5dump extracted from the same files)."""
    path = os.path.join(REF_TESTS, name, "FCIDUMP_chol")
    if not os.path.exists(path):
        pytest.skip("reference test files not present")
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    if "hcore_mod" not in z.files:   # unrestricted integrals (utils/integral.cpp:664-741)
        d = read_fcidump_chol(path, "u")
        assert [d.nelec, d.norb, d.ms, d.nchol] == [int(v) for v in z["header"]]
        for got, key in ((d.h1, "hcore_up"), (d.h1_dn, "hcore_dn"), (d.h1_mod, "hcore_mod_up"),
                         (d.h1_mod_dn, "hcore_mod_dn"), (d.chol, "chol_up"), (d.chol_dn, "chol_dn")):
            assert np.array_equal(got.ravel(), np.asarray(z[key]).ravel()), key
        assert d.ecore == float(np.asarray(z["energy_core"]).ravel()[0])
        return
    d = read_fcidump_chol(path)
    assert [d.nelec, d.norb, d.ms, d.nchol] == [int(v) for v in z["header"]]
    assert np.array_equal(d.h1.ravel(), np.asarray(z["hcore"]).ravel())
    assert np.array_equal(d.h1_mod.ravel(), np.asarray(z["hcore_mod"]).ravel())
    assert np.array_equal(d.chol.ravel(), np.asarray(z["chol"]).ravel())
    assert d.ecore == float(np.asarray(z["energy_core"]).ravel()[0])


# ---------------------------------------------------------------- afqmc.json
from vmc_amd.dqmc import read_afqmc_json  # noqa: E402


def test_afqmc_json_keys_and_defaults(tmp_path):
    """utils/input.cpp:61-276: the keys the phaseless walk reads, case-insensitive
    (boost iptree), with the reference's defaults where absent."""
    p = tmp_path / "afqmc.json"
    p.write_text(json.dumps({"System": {"integrals": "FCIDUMP_chol"},
                             "wavefunction": {"left": "UHF", "right": "rhf"},
                             "sampling": {"seed": 4321, "Phaseless": True, "dt": 0.005, "nsteps": 50,
                                          "nwalk": 20, "stochasticIter": 50, "choleskyThreshold": 0.001,
                                          "orthoSteps": 20},
                             "print": {"writeOneRDM": True}}))
    s = read_afqmc_json(str(p))
    assert (s.integrals, s.int_type, s.left, s.right) == ("FCIDUMP_chol", "r", "uhf", "rhf")
    assert (s.seed, s.phaseless, s.dt, s.nsteps, s.nwalk) == (4321, True, 0.005, 50, 20)
    assert (s.stochastic_iter, s.cholesky_threshold, s.ortho_steps) == (50, 0.001, 20)
    assert (s.burn_iter, s.ene0_guess, s.weight_cap, s.write_one_rdm, s.scratch_dir) == (0, 1e10, -1.0, True, "./")
    p.write_text(json.dumps({"sampling": {"seed": 1}}))
    d = read_afqmc_json(str(p))
    assert (d.integrals, d.dt, d.nsteps, d.nwalk, d.stochastic_iter, d.ortho_steps, d.cholesky_threshold,
            d.phaseless) == ("FCIDUMP", 0.1, 10, 100, 10000, 50, 0.005, False)


# ---------------------------------------------------------------- spinRDM.0.0.txt
from vmc_amd.formats import read_spin_rdm  # noqa: E402


def test_spin_rdm_of_a_determinant(tmp_path):
    """For one determinant with spin-orbital density D (idempotent, trace n) the
    2-RDM is Gamma_abcd = D_ac D_bd - D_ad D_bc, and its contraction over the second
    index divided by n - 1 is D again (closed form); listing only a < b, c < d, as
    the file does, must give the same.  Mixed spin blocks of D vanish for an
    unrestricted determinant, and the reader returns the even/odd blocks."""
    rng = np.random.default_rng(6)
    N, na, nb = 5, 3, 2
    Cu = np.linalg.qr(rng.standard_normal((N, na)))[0]
    Cd = np.linalg.qr(rng.standard_normal((N, nb)))[0]
    Du, Dd = Cu @ Cu.T, Cd @ Cd.T
    D = np.zeros((2 * N, 2 * N))
    D[0::2, 0::2], D[1::2, 1::2] = Du, Dd
    lines = []
    for a in range(2 * N):
        for b in range(a + 1, 2 * N):
            for c in range(2 * N):
                for d in range(c + 1, 2 * N):
                    g = D[a, c] * D[b, d] - D[a, d] * D[b, c]
                    if abs(g) > 0:
                        lines.append("%d %d %d %d %.17g" % (a, b, c, d, g))
    p = tmp_path / "spinRDM.0.0.txt"
    p.write_text("\n".join(lines) + "\n")
    gu, gd = read_spin_rdm(str(p), N, na + nb)
    assert np.allclose(gu, Du, atol=1e-13) and np.allclose(gd, Dd, atol=1e-13)
