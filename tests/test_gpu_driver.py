"""GPU driver (calcMixedEstimatorLongProp over the C ABI) end to end.

Replay mode draws the reference's libstdc++ stream on the host, so:
  * in FP64 the block weights/energies match the oracle's driver fed the same
    stream (1e-9: the comb turns 1e-12 weight differences into identical
    resampling decisions, so trajectories stay locked);
  * in ref_compat mode with 4 virtual ranks the GPU reproduces the reference's
    own tests/DQMC/*/samples.ref (the reference's pass tolerance is 1e-7).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from vmc_amd.afqmc import AFQ_RNG_DEVICE, run_long_prop
from vmc_amd.formats import load_case

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["rhf_rhf", "uhf_uhf", "rhf_uhf", "uhf_rhf", "uhf_uhf_ui"]   # uhf_uhf_ui: unrestricted integrals


def kw(c, **extra):
    d = dict(seed=c.seed, dt=c.dt, nsteps=c.nsteps, nsweeps=c.nsweeps, ortho_steps=c.ortho_steps,
             nwalk=c.nwalk, nranks=c.nranks, walker_rhf=c.walker_rhf,
             cholesky_threshold=c.cholesky_threshold)
    d.update(extra)
    return d


@pytest.mark.parametrize("name", CASES)
def test_driver_matches_oracle_driver_fp64(name):
    c = load_case(os.path.join(GOLD, name + ".npz"))
    nb = 8
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn,
                              **kw(c, max_steps=nb * c.nsteps - 1))
    rc, wo, eo, reso = O.run_long_prop(c.ham, c.ham.trial_rhf, c.ref_up, c.ref_dn, c.init_up,
                                       c.init_dn, **kw(c, max_steps=nb * c.nsteps - 1))
    assert rc == 0
    assert res.nchol_ene == reso.nchol_ene
    assert abs(res.delta - reso.delta) < 1e-12
    assert abs(res.ref_energy - c.samples[0, 1]) < 1e-8 * abs(c.samples[0, 1]) + abs(res.delta)
    assert np.max(np.abs(e[:nb] - eo[:nb]) / np.abs(eo[:nb])) < 1e-9
    assert np.max(np.abs(w[:nb] - wo[:nb]) / wo[:nb]) < 1e-9


@pytest.mark.parametrize("name", CASES)
def test_driver_reproduces_samples_ref(name):
    c = load_case(os.path.join(GOLD, name + ".npz"))
    nb = 10
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn,
                              **kw(c, ref_compat=True, max_steps=nb * c.nsteps - 1))
    assert np.all(np.abs(e[:nb] - c.samples[:nb, 1]) < 1e-7), np.abs(e[:nb] - c.samples[:nb, 1])
    assert np.all(np.abs(w[:nb] - c.samples[:nb, 0]) < 1e-6 * c.samples[:nb, 0])


def test_driver_full_run_last_line():
    """The reference's own pass criterion: last samples.dat line within 1e-7."""
    c = load_case(os.path.join(GOLD, "rhf_rhf.npz"))
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn, **kw(c, ref_compat=True))
    assert abs(e[-1] - c.samples[-1, 1]) < 1e-7
    assert abs(w[-1] - c.samples[-1, 0]) < 1e-5


def test_driver_device_rng_statistics():
    """Device RNG: a different stream, same physics -- block energies agree with
    samples.ref statistically (blocked means within a few error bars)."""
    c = load_case(os.path.join(GOLD, "uhf_rhf.npz"))
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn,
                              **kw(c, rng_mode=AFQ_RNG_DEVICE, nwalk=80, nranks=4))
    m1, e1 = O.blocking(w[10:], e[10:])
    m2, e2 = O.blocking(c.samples[10:, 0], c.samples[10:, 1])
    err = np.hypot(max(e1, 1e-3), max(e2, 1e-3))
    assert abs(m1 - m2) < 4 * err, (m1, e1, m2, e2)
