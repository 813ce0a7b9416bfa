"""Unrestricted integrals (intType "u", propagatePhaselessU, DQMC/DQMCWalker.cpp:106-187,
506-554) on the GPU beyond the shared parity tables: the whole reference run of
tests/DQMC/uhf_uhf_ui against its samples.ref, what the path does not take, and the
reduced-precision exchange on it (the per-spin rotated Cholesky rows feed it as they
feed the FP64 kernels)."""
import os

import numpy as np
import pytest

from test_gpu_parity import GOLD, build, perturbed_walkers, rel
from vmc_amd.abi import AFQ_PREC_BF16, AFQ_PREC_BF16X3, AFQ_PREC_FP64
from vmc_amd.afqmc import AfqError, Context, run_long_prop
from vmc_amd.formats import load_case

pytestmark = pytest.mark.gpu


def test_full_run_reproduces_samples_ref():
    """The reference's own pass criterion (tests/runDQMC.sh:9): every samples.dat line
    within 1e-7 of samples.ref, 4 virtual ranks, the reference's FP32 VHS."""
    c = load_case(os.path.join(GOLD, "uhf_uhf_ui.npz"))
    w, e, res = run_long_prop(c.ham, c.ref_up, c.ref_dn, c.init_up, c.init_dn, seed=c.seed, dt=c.dt,
                              nsteps=c.nsteps, nsweeps=c.nsweeps, ortho_steps=c.ortho_steps, nwalk=c.nwalk,
                              nranks=c.nranks, walker_rhf=c.walker_rhf, ref_compat=True,
                              cholesky_threshold=c.cholesky_threshold)
    assert np.max(np.abs(e - c.samples[:, 1])) < 1e-7, np.max(np.abs(e - c.samples[:, 1]))
    assert np.max(np.abs(w - c.samples[:, 0]) / c.samples[:, 0]) < 1e-7


def test_unsupported_combinations_fail_loudly():
    h, wrhf, ctx, orc = build("u_33_31", 2)
    with pytest.raises(AfqError):
        ctx.set_precision(prec_vhs=AFQ_PREC_BF16X3)
    with pytest.raises(AfqError):
        ctx.set_prop_precision(AFQ_PREC_BF16X3)
    ctx.close()
    with pytest.raises(AfqError):
        Context(h, 2, walker_rhf=True)
    with pytest.raises(AfqError):
        Context(h, 2, walker_rhf=False, prec_vhs=AFQ_PREC_BF16X3)


@pytest.mark.parametrize("prec,tol", [(AFQ_PREC_BF16X3, 5e-6), (AFQ_PREC_BF16, 5e-2)])
def test_reduced_precision_exchange(prec, tol):
    W = 6
    h, wrhf, ctx, orc = build("u_n60", W)
    dets = perturbed_walkers(h, W, wrhf, np.random.default_rng(9), 0.1)
    ctx.set_walkers(dets)
    e64, _ = ctx.measure()
    ctx.set_precision(prec_exchange=prec)
    elp = ctx.measure(want_ho=False)
    ctx.set_precision(prec_exchange=AFQ_PREC_FP64)
    assert rel(elp, e64) < tol, rel(elp, e64)
    ctx.close()
