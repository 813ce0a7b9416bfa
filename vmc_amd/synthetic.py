"""Seeded synthetic Hamiltonians in the reference's FCIDUMP_chol fields.

Generator (same for every workload):
  h   = sym(A)/N + diag(linspace(-2, 0, N)),   A_ij ~ N(0,1), sym(A) = (A + A^T)/2
  L_g = symmetric, upper triangle iid N(0, 0.3^2)/sqrt(N), mirrored, so that the
        ERI (ij|kl) = sum_g L_g,ij L_g,kl is positive semidefinite by construction
  hcore_mod = h - 1/2 sum_g L_g L_g   (scripts/QMCUtils.py:644-645)
  ecore = 0
Trial: lowest-n eigenvectors of the mean-field Fock matrix F = h + J - K built from
the core-guess density (RHF), or a short UHF iteration started from that RHF with
a small alpha/beta HOMO-LUMO rotation (UHF).

Workloads (names used by tests and bench.py):
  c0  N=10,  5a+5b,   G=40,   RHF trial, 128 walkers   (reference-CPU-runnable case)
  c1  N=100, 20a+20b, G=400,  UHF trial, 1024 walkers
  c3  the c1 system with 4096 walkers (surrogate control variate)
  c4  N=300, 60a+60b, G=1200, UHF trial, 4096 walkers
  c2  64 systems, N ~ U{20..120}, n = round(N/5), G = 4N, 256 walkers each
"""
import numpy as np

from .abi import HamArrays

DEFAULT_SEED = 20261202


def _eigh_lowest(F, n):
    w, v = np.linalg.eigh(F)
    return np.ascontiguousarray(v[:, :n])


def _jk(L, C):
    """Coulomb and exchange of the density C C^T with ERIs from L (G x N x N)."""
    G, N, _ = L.shape
    D = C @ C.T
    Lf = L.reshape(G, N * N)
    J = ((Lf @ D.ravel()) @ Lf).reshape(N, N)
    LC = (L.reshape(G * N, N) @ C).reshape(G, N, -1)   # rows (g, i) of L_g C
    LCt = LC.transpose(1, 0, 2).reshape(N, -1)           # i x (g, n)
    K = LCt @ LCt.T
    return J, K


def make_system(norb, nalpha, nbeta, nchol, seed=DEFAULT_SEED, trial="rhf",
                uhf_iters=8, uhf_kick=0.15):
    rng = np.random.default_rng(seed)
    N = norb
    A = rng.standard_normal((N, N))
    h = 0.5 * (A + A.T) / N + np.diag(np.linspace(-2.0, 0.0, N))
    L = np.empty((nchol, N, N))
    iu = np.triu_indices(N)
    scale = 0.3 / np.sqrt(N)
    for g in range(nchol):
        u = rng.standard_normal(len(iu[0])) * scale
        M = np.zeros((N, N))
        M[iu] = u
        L[g] = M + np.triu(M, 1).T
    M = L.reshape(nchol * N, N)          # sum_g L_g L_g = M^T M for symmetric L_g
    h_mod = h - 0.5 * (M.T @ M)
    # RHF from the core guess
    C0 = _eigh_lowest(h, nalpha)
    J, K = _jk(L, C0)
    F = h + 2.0 * J - K
    Ca = _eigh_lowest(F, nalpha)
    Cb = _eigh_lowest(F, nbeta)
    if trial == "uhf":
        # break spin symmetry with opposite HOMO-LUMO rotations, then iterate
        w, v = np.linalg.eigh(F)
        def rot(n, sgn):
            C = v[:, :n].copy()
            if n < N:
                c, s = np.cos(uhf_kick), np.sin(sgn * uhf_kick)
                homo, lumo = v[:, n - 1].copy(), v[:, n].copy()
                C[:, n - 1] = c * homo + s * lumo
            return C
        Ca, Cb = rot(nalpha, 1.0), rot(nbeta, -1.0)
        for _ in range(uhf_iters):
            Ja, Ka = _jk(L, Ca)
            Jb, Kb = _jk(L, Cb)
            Fa = h + Ja + Jb - Ka
            Fb = h + Ja + Jb - Kb
            Ca, Cb = _eigh_lowest(Fa, nalpha), _eigh_lowest(Fb, nbeta)
        return HamArrays(N, nalpha, nbeta, 0.0, h, h_mod, L, Ca, Cb, trial_rhf=False)
    return HamArrays(N, nalpha, nbeta, 0.0, h, h_mod, L, Ca, Ca[:, :nbeta], trial_rhf=True)


WORKLOADS = {
    "c0": dict(norb=10, nalpha=5, nbeta=5, nchol=40, trial="rhf", nwalk=128, dt=0.005,
               steps=200, ortho=5),
    "c1": dict(norb=100, nalpha=20, nbeta=20, nchol=400, trial="uhf", nwalk=1024, dt=0.005,
               steps=200, ortho=5),
    "c3": dict(norb=100, nalpha=20, nbeta=20, nchol=400, trial="uhf", nwalk=4096, dt=0.005,
               steps=200, ortho=5),
    "c4": dict(norb=300, nalpha=60, nbeta=60, nchol=1200, trial="uhf", nwalk=4096, dt=0.005,
               steps=200, ortho=5),
}


def workload(name, seed=DEFAULT_SEED):
    """(HamArrays, spec dict) of a named workload; walker_rhf follows the reference's
    selection rule (RHF walker iff nalpha == nbeta and an RHF right state)."""
    spec = dict(WORKLOADS[name])
    ham = make_system(spec["norb"], spec["nalpha"], spec["nbeta"], spec["nchol"], seed=seed,
                      trial=spec["trial"])
    spec["walker_rhf"] = spec["trial"] == "rhf"
    return ham, spec


def batched_systems(nsys=64, seed=DEFAULT_SEED, nmin=20, nmax=120):
    """The uneven multi-system workload: N ~ U{nmin..nmax}, n = round(N/5), G = 4N."""
    rng = np.random.default_rng(seed + 7)
    out = []
    for s in range(nsys):
        N = int(rng.integers(nmin, nmax + 1))
        n = max(1, int(round(N / 5)))
        out.append(make_system(N, n, n, 4 * N, seed=seed + 100 + s, trial="uhf"))
    return out
