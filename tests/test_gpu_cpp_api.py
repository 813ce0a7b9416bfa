"""The C++ mirror of the reference interfaces (include/vmc_afqmc.hpp) drives the
population like calcMixedEstimatorLongProp drives DQMCWalker objects; its final
walkers, overlaps, weights, hamAndOverlap, force bias and local energies match
the CPU oracle fed the same draws (1e-10 relative).  The host program is built
by __graft_entry__.build() (g++, tests/cpp/api_parity.cpp)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import SYSTEMS, TOL, perturbed_walkers, rel, rel_each

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_parity")


@pytest.mark.parametrize("name", ["uhf_uhf", "rhf_rhf", "uhf_33_31"])
def test_cpp_api_matches_oracle(name, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/api_parity is not built (run __graft_entry__.build())"
    obj, wrhf = SYSTEMS[name]
    obj = obj()
    if hasattr(obj, "ham"):
        h, wrhf, ru, rd = obj.ham, obj.walker_rhf, obj.ref_up, obj.ref_dn
    else:
        h, ru, rd = obj, obj.phi_up, obj.phi_dn
    trhf = h.trial_rhf
    W, steps, dt = 5, 3, 0.005
    rng = np.random.default_rng(17)
    dets = perturbed_walkers(h, W, wrhf, rng, 0.05)
    draws = rng.standard_normal((steps, W, h.nchol))
    orc = O.Oracle(h, W, wrhf)
    orc.prep_prop(ru, rd, dt)
    orc.set_walkers(dets)
    o = orc.ham_and_overlap_dets(h.serial_det(ru, rd)[None, :], h.nchol)[0]
    eshift = float((o[0] / o[1]).real)
    N, na, nb, G = h.norb, h.nalpha, h.nbeta, h.nchol
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([N, na, nb, G, W, steps, int(wrhf), int(trhf)], np.int32).tofile(f)
        np.array([h.ecore, dt, eshift], np.float64).tofile(f)
        for a in (h.h1, h.h1_mod, h.chol, h.phi_up, h.phi_dn, ru, rd):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(dets, np.complex128).tofile(f)
        np.ascontiguousarray(draws, np.float64).tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for s in range(steps):
        orc.propagate(draws[s], eshift, 100.0)
        if s == 1:
            orc.orthogonalize()
    ms = N * (na + nb)
    buf = np.fromfile(out, np.uint8)
    off = 0

    def take(dtype, n):
        nonlocal off
        a = buf[off:off + n * np.dtype(dtype).itemsize].view(dtype)
        off += n * np.dtype(dtype).itemsize
        return a

    dg = take(np.complex128, W * ms).reshape(W, ms)
    og = take(np.complex128, W)
    wg = take(np.float64, W)
    hog = take(np.complex128, 2 * W).reshape(W, 2)
    fbg = take(np.complex128, W * G).reshape(W, G)
    elg = take(np.float64, W)
    assert off == buf.size
    do, oo, wo = orc.get_walkers()
    live = wo > 0
    assert live.any()
    # a walker of weight zero is skipped by the re-orthonormalisation and the force bias
    assert rel(dg[live], do[live]) < TOL and rel(wg, wo) < TOL
    assert rel_each(og[live], oo[live]) < TOL
    for w in range(W):
        if live[w]:
            assert rel(fbg[w], orc.force_bias(w)) < TOL
            ho = orc.ham_and_overlap(w)
            assert rel_each(hog[w], ho) < TOL
            assert abs(elg[w] - (ho[0] / ho[1]).real) < TOL * abs(elg[w])
