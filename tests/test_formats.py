"""Reference output formats (CPU): the 1-RDM files match what std::ostream writes
for the same numbers (writeOneRDM, DQMC/MixedEstimator.cpp:52-112: default
stream format, entries followed by two spaces), and samples.dat matches
boost::format("%.7e      %.10e \\n") (DQMC/MixedEstimator.cpp:681), which is C's printf."""
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
