"""GPU mixed-estimator one-body density matrix (DQMC/MixedEstimator.cpp:492-510)
against the CPU oracle.

afq_rdm_accumulate sums w Re(theta) of every walker of nonzero weight on the
device; afq_get_one_rdm forms Phi_s (sum w Re theta_s)^T / sum w.  Parity with the
oracle's running average is to 1e-10 relative (FP64 throughout; the oracle
updates a running average, the device a plain sum, so they differ by rounding
only).  Shapes: the four reference fixtures (RHF-RHF is one slot of
multiplicity 2), a synthetic nalpha != nbeta system and the N=100 system, with
walker counts below and above the accumulation's 32 walker chunks.  The driver
test compares whole runs of the reference fixtures (uhf_uhf_ui with unrestricted
integrals, where the reference writes the per-spin pair as here), where the 1-RDM is
accumulated once step dt > 10; both sides run the reference's FP32 VHS
(ref_compat), where whole trajectories agree to the reference's own test
tolerance, 1e-7 (tests/runDQMC.sh:9), and so does the 1-RDM.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import GOLD, build, perturbed_walkers, rel
from vmc_amd.afqmc import run_long_prop
from vmc_amd.formats import load_case

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize("name,W", [("rhf_rhf", 9), ("uhf_uhf", 9), ("rhf_uhf", 9), ("uhf_rhf", 9),
                                    ("uhf_33_31", 40), ("c1_small", 300)])
def test_one_rdm_matches_oracle(name, W):
    h, wrhf, ctx, orc = build(name, W)
    rng = np.random.default_rng(21)
    ctx.rdm_reset()
    state = {1: (None, 0.0), 0: (None, 0.0)}
    for m in range(2):  # two measurements continue one average
        dets = perturbed_walkers(h, W, wrhf, rng)
        wts = rng.uniform(0.2, 2.0, W)
        wts[(3 + m) % W] = 0.0
        ctx.set_walkers(dets, weights=wts)
        orc.set_walkers(dets, weights=wts)
        ctx.rdm_accumulate()
        for summed in (1, 0):
            state[summed] = orc.one_rdm(summed, *state[summed])
            rg, cwg = ctx.one_rdm(summed)
            ro, cwo = state[summed]
            assert abs(cwg - cwo) < 1e-13 * cwo, (name, m, summed)
            assert rel(rg, ro) < TOL, (name, m, summed, rel(rg, ro))
    per, _ = ctx.one_rdm(0)
    assert abs(np.trace(per[0]) - h.nalpha) < 1e-10 and abs(np.trace(per[1]) - h.nbeta) < 1e-10
    assert np.allclose(per[0] @ h.phi_up, h.phi_up, atol=1e-10)


def test_one_rdm_empty_and_reset():
    h, wrhf, ctx, orc = build("uhf_uhf", 4)
    ctx.set_walkers(perturbed_walkers(h, 4, wrhf, np.random.default_rng(2)), weights=np.zeros(4))
    ctx.rdm_reset()
    r, cw = ctx.one_rdm(0)
    assert cw == 0.0 and not r.any()
    ctx.rdm_accumulate()  # every weight zero: nothing is added
    r, cw = ctx.one_rdm(1)
    assert cw == 0.0 and not r.any()


@pytest.mark.parametrize("name", ["rhf_rhf", "uhf_uhf", "uhf_uhf_ui"])
def test_driver_one_rdm_matches_oracle_driver(name):
    c = load_case(os.path.join(GOLD, name + ".npz"))
    N = c.ham.norb
    shape = (N, N) if c.ham.trial_rhf else (2, N, N)
    rg, ro = np.zeros(shape), np.zeros(shape)
    kw = dict(seed=c.seed, dt=c.dt, nsteps=c.nsteps, nsweeps=c.nsweeps, ortho_steps=c.ortho_steps,
              nwalk=c.nwalk, nranks=c.nranks, walker_rhf=c.walker_rhf,
              cholesky_threshold=c.cholesky_threshold, ref_compat=True)
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn, one_rdm=rg, **kw)
    rc, wo, eo, reso = O.run_long_prop(c.ham, c.ham.trial_rhf, c.ref_up, c.ref_dn, c.init_up, c.init_dn,
                                       one_rdm=ro, **kw)
    assert rc == 0
    assert res.one_rdm_weight > 0
    assert abs(res.one_rdm_weight - reso.one_rdm_weight) < 1e-9 * reso.one_rdm_weight
    assert rel(rg, ro) < 1e-7, rel(rg, ro)
    ne = c.ham.nalpha + c.ham.nbeta
    tr = np.trace(rg) if rg.ndim == 2 else np.trace(rg[0]) + np.trace(rg[1])
    assert abs(tr - ne) < 1e-9
