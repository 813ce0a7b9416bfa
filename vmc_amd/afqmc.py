"""ctypes binding of the HIP library (vmc_amd/libvmc_afqmc.so, C ABI in
include/vmc_afqmc.h).

There is no CPU fallback: if the library is missing or no device is present the
calls raise.  Build it with `python -m vmc_amd.build` (or __graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

from .abi import (AFQ_ENERGY_EXACT, AFQ_PREC_FP64, AfqHam, AfqParams, AfqTrial, DP, dptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvmc_afqmc.so")

_lib = None


class AfqError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfqError("HIP library %s not built (run python -m vmc_amd.build)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        i32, i64, u64, d = C.c_int32, C.c_int64, C.c_uint64, C.c_double
        L.afq_version.restype = C.c_int
        L.afq_last_error.restype = C.c_char_p
        L.afq_last_error.argtypes = [vp]
        L.afq_create.argtypes = [C.POINTER(AfqHam), C.POINTER(AfqTrial), C.POINTER(AfqParams), C.POINTER(vp)]
        L.afq_destroy.argtypes = [vp]
        L.afq_set_nchol_ene.argtypes = [vp, i32]
        L.afq_prep_prop.argtypes = [vp, DP, DP, DP, DP, d, d]
        L.afq_set_walkers.argtypes = [vp, DP, DP, DP]
        L.afq_get_walkers.argtypes = [vp, DP, DP, DP]
        L.afq_get_weights.argtypes = [vp, DP]
        L.afq_set_weights.argtypes = [vp, DP]
        L.afq_propagate.argtypes = [vp, DP, u64, u64, d, DP, C.POINTER(i64)]
        L.afq_orthogonalize.argtypes = [vp]
        L.afq_measure.argtypes = [vp, i32, d, DP, DP]
        L.afq_ham_and_overlap.argtypes = [vp, DP, i32, i32, DP]
        L.afq_force_bias.argtypes = [vp, DP]
        L.afq_overlap.argtypes = [vp, DP]
        L.afq_comb.argtypes = [vp, d, DP]
        L.afq_stage_times.argtypes = [vp, DP, i32]
        L.afq_sync.argtypes = [vp]
        L.afq_debug_gemm_f64.argtypes = [DP, DP, DP, i32, i32, i32]
        L.afq_debug_copy.argtypes = [i64]
        L.afq_debug_gemm_bf16.argtypes = [DP, DP, C.POINTER(C.c_float), i32, i32, i32, i32]
        _lib = L
    return _lib


STAGES = ["force_bias", "fields", "vhs", "propagate", "green", "weight", "ortho", "energy", "comb"]


def c2f(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


def debug_gemm_f64(A, B):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.zeros((M, N))
    rc = lib().afq_debug_gemm_f64(dptr(A), dptr(B), dptr(Cm), M, N, K)
    if rc != 0:
        raise AfqError("debug_gemm: %s" % lib().afq_last_error(None).decode())
    return Cm


def debug_gemm_bf16(A, B, passes=3):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    Cm = np.zeros((M, N), np.float32)
    rc = lib().afq_debug_gemm_bf16(dptr(A), dptr(B), Cm.ctypes.data_as(C.POINTER(C.c_float)), M, N, K, passes)
    if rc != 0:
        raise AfqError("debug_gemm_bf16: %s" % lib().afq_last_error(None).decode())
    return Cm


def debug_copy(nbytes):
    """Run the plain copy calibration kernel over nbytes (read) + nbytes (write)."""
    if lib().afq_debug_copy(int(nbytes)) != 0:
        raise AfqError("debug_copy: %s" % lib().afq_last_error(None).decode())


class Context:
    """One afq_ctx: a rank's walker population on one GPU."""

    def __init__(self, ham, nwalk, walker_rhf, ref_compat=False, prec_fb=AFQ_PREC_FP64,
                 prec_vhs=AFQ_PREC_FP64, prec_exchange=AFQ_PREC_FP64, device=0, weight_cap=0.0):
        self.h = ham
        self.nwalk = int(nwalk)
        self._hs = ham.ham_struct()
        self._ts = ham.trial_struct()
        self._ps = AfqParams(self.nwalk, int(walker_rhf), int(ref_compat), prec_fb, prec_vhs,
                             prec_exchange, device, weight_cap)
        p = C.c_void_p()
        rc = lib().afq_create(C.byref(self._hs), C.byref(self._ts), C.byref(self._ps), C.byref(p))
        if rc != 0:
            raise AfqError("afq_create(%d): %s" % (rc, lib().afq_last_error(None).decode()))
        self.ptr = p

    def _chk(self, rc, what):
        if rc != 0:
            raise AfqError("%s(%d): %s" % (what, rc, lib().afq_last_error(self.ptr).decode()))

    def close(self):
        if getattr(self, "ptr", None):
            lib().afq_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_nchol_ene(self, n):
        self._chk(lib().afq_set_nchol_ene(self.ptr, int(n)), "afq_set_nchol_ene")

    def prep_prop(self, ref_up, ref_dn, dt, ene0=0.0, rdm_up=None, rdm_dn=None):
        ru = np.ascontiguousarray(ref_up, dtype=np.float64)
        rd = np.ascontiguousarray(ref_dn, dtype=np.float64)
        a = None if rdm_up is None else np.ascontiguousarray(rdm_up, dtype=np.float64)
        b = None if rdm_dn is None else np.ascontiguousarray(rdm_dn, dtype=np.float64)
        self._chk(lib().afq_prep_prop(self.ptr, dptr(ru), dptr(rd), dptr(a), dptr(b), dt, ene0),
                  "afq_prep_prop")

    def set_walkers(self, dets, overlaps=None, weights=None):
        d = c2f(dets)
        o = None if overlaps is None else c2f(overlaps)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        self._chk(lib().afq_set_walkers(self.ptr, dptr(d), dptr(o), dptr(w)), "afq_set_walkers")

    def get_walkers(self):
        d = np.zeros((self.nwalk, self.h.mat_size), np.complex128)
        o = np.zeros(self.nwalk, np.complex128)
        w = np.zeros(self.nwalk)
        self._chk(lib().afq_get_walkers(self.ptr, dptr(d.view(np.float64)), dptr(o.view(np.float64)),
                                        dptr(w)), "afq_get_walkers")
        return d, o, w

    def weights(self):
        w = np.zeros(self.nwalk)
        self._chk(lib().afq_get_weights(self.ptr, dptr(w)), "afq_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._chk(lib().afq_set_weights(self.ptr, dptr(w)), "afq_set_weights")

    def propagate(self, draws=None, eshift=0.0, seed=0, step=0, want_factors=True):
        dr = None if draws is None else np.ascontiguousarray(draws, dtype=np.float64)
        f = np.zeros(self.nwalk) if want_factors else None
        nc = C.c_int64(0)
        self._chk(lib().afq_propagate(self.ptr, dptr(dr), seed, step, eshift, dptr(f), C.byref(nc)),
                  "afq_propagate")
        return f, nc.value

    def orthogonalize(self):
        self._chk(lib().afq_orthogonalize(self.ptr), "afq_orthogonalize")

    def measure(self, delta=0.0, mode=AFQ_ENERGY_EXACT, want_ho=True):
        e = np.zeros(self.nwalk)
        ho = np.zeros((self.nwalk, 2), np.complex128) if want_ho else None
        self._chk(lib().afq_measure(self.ptr, mode, delta, dptr(e),
                                    dptr(None if ho is None else ho.view(np.float64))), "afq_measure")
        return (e, ho) if want_ho else e

    def ham_and_overlap(self, dets, nchol_ene):
        dets = np.atleast_2d(np.asarray(dets, dtype=np.complex128))
        out = np.zeros((dets.shape[0], 2), np.complex128)
        self._chk(lib().afq_ham_and_overlap(self.ptr, dptr(c2f(dets)), dets.shape[0], int(nchol_ene),
                                            dptr(out.view(np.float64))), "afq_ham_and_overlap")
        return out

    def force_bias(self):
        fb = np.zeros((self.nwalk, self.h.nchol), np.complex128)
        self._chk(lib().afq_force_bias(self.ptr, dptr(fb.view(np.float64))), "afq_force_bias")
        return fb

    def overlaps(self):
        o = np.zeros(self.nwalk, np.complex128)
        self._chk(lib().afq_overlap(self.ptr, dptr(o.view(np.float64))), "afq_overlap")
        return o

    def comb(self, zeta):
        t = np.zeros(1)
        self._chk(lib().afq_comb(self.ptr, zeta, dptr(t)), "afq_comb")
        return t[0]

    def stage_times(self, reset=True):
        ms = np.zeros(9)
        self._chk(lib().afq_stage_times(self.ptr, dptr(ms), int(reset)), "afq_stage_times")
        return dict(zip(STAGES, ms))

    def sync(self):
        self._chk(lib().afq_sync(self.ptr), "afq_sync")


AFQ_RNG_REPLAY = 0
AFQ_RNG_DEVICE = 1


class RunParams(C.Structure):
    _fields_ = [
        ("seed", C.c_int64), ("dt", C.c_double), ("nsteps", C.c_int32), ("nsweeps", C.c_int32),
        ("ortho_steps", C.c_int32), ("burn_iter", C.c_int32), ("nwalk", C.c_int32),
        ("nranks", C.c_int32), ("walker_rhf", C.c_int32), ("ref_compat", C.c_int32),
        ("cholesky_threshold", C.c_double), ("ene0_guess", C.c_double), ("weight_cap", C.c_double),
        ("max_steps", C.c_int32), ("rng_mode", C.c_int32), ("warmup_steps", C.c_int32),
        ("prec_fb", C.c_int32), ("prec_vhs", C.c_int32), ("prec_exchange", C.c_int32),
        ("energy_mode", C.c_int32), ("device", C.c_int32),
    ]


class RunResult(C.Structure):
    _fields_ = [("ref_energy", C.c_double), ("nchol_ene", C.c_int32), ("delta", C.c_double),
                ("n_large_deviations", C.c_int64), ("prop_time", C.c_double),
                ("ene_time", C.c_double), ("timed_steps", C.c_int64), ("timed_wall_s", C.c_double),
                ("stage_ms", C.c_double * 9)]


def run_long_prop(ham, ref_up, ref_dn, init_up, init_dn, *, seed, dt, nsteps, nsweeps, ortho_steps,
                  nwalk, nranks=1, walker_rhf, ref_compat=False, cholesky_threshold=1e-3,
                  burn_iter=0, ene0_guess=1e10, weight_cap=0.0, max_steps=0,
                  rng_mode=AFQ_RNG_REPLAY, warmup_steps=0, prec_fb=AFQ_PREC_FP64,
                  prec_vhs=AFQ_PREC_FP64, prec_exchange=AFQ_PREC_FP64,
                  energy_mode=AFQ_ENERGY_EXACT, device=0):
    """calcMixedEstimatorLongProp on the GPU; returns (weights, energies, RunResult)."""
    L = lib()
    if not hasattr(L, "_run_bound"):
        L.afq_run_long_prop.argtypes = [C.POINTER(AfqHam), C.POINTER(AfqTrial), DP, DP, DP, DP,
                                        C.POINTER(RunParams), DP, DP, C.POINTER(RunResult)]
        L._run_bound = True
    hs, ts = ham.ham_struct(), ham.trial_struct()
    p = RunParams(seed, dt, nsteps, nsweeps, ortho_steps, burn_iter, nwalk, nranks, int(walker_rhf),
                  int(ref_compat), cholesky_threshold, ene0_guess, weight_cap, max_steps, rng_mode,
                  warmup_steps, prec_fb, prec_vhs, prec_exchange, energy_mode, device)
    w = np.zeros(nsweeps)
    e = np.zeros(nsweeps)
    res = RunResult()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ref_up, ref_dn, init_up, init_dn)]
    rc = L.afq_run_long_prop(C.byref(hs), C.byref(ts), *[dptr(a) for a in arrs], C.byref(p),
                             dptr(w), dptr(e), C.byref(res))
    if rc != 0:
        raise AfqError("afq_run_long_prop(%d): %s" % (rc, L.afq_last_error(None).decode()))
    return w, e, res
