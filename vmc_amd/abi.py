"""ctypes mirror of the structs in include/vmc_afqmc.h.

The same structs feed the HIP library (vmc_amd/libvmc_afqmc.so) and, in the
tests, the CPU oracle, so both sides always see byte-identical inputs.
"""
import ctypes as C

import numpy as np

AFQ_OK = 0
AFQ_PREC_FP64 = 0
AFQ_PREC_FP32 = 1
AFQ_PREC_BF16X3 = 2
AFQ_PREC_BF16 = 3
AFQ_PREC_FP8 = 4
AFQ_ENERGY_EXACT = 0
AFQ_ENERGY_SURROGATE = 1
AFQ_SURR_MLP = 0
AFQ_SURR_EXCHANGE = 1

DP = C.POINTER(C.c_double)


class AfqHam(C.Structure):
    _fields_ = [
        ("norb", C.c_int32),
        ("nalpha", C.c_int32),
        ("nbeta", C.c_int32),
        ("nchol", C.c_int32),
        ("ecore", C.c_double),
        ("h1", DP),
        ("h1_mod", DP),
        ("chol", DP),
        ("h1_dn", DP),
        ("h1_mod_dn", DP),
        ("chol_dn", DP),
    ]


class AfqTrial(C.Structure):
    _fields_ = [("rhf", C.c_int32), ("phi_up", DP), ("phi_dn", DP)]


class AfqParams(C.Structure):
    _fields_ = [
        ("nwalk", C.c_int32),
        ("walker_rhf", C.c_int32),
        ("ref_compat", C.c_int32),
        ("prec_fb", C.c_int32),
        ("prec_vhs", C.c_int32),
        ("prec_exchange", C.c_int32),
        ("device", C.c_int32),
        ("weight_cap", C.c_double),
    ]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return DP()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(DP)


def cplx_view(a):
    """complex128 array -> interleaved float64 view (no copy)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


class HamArrays:
    """Owns the numpy buffers behind one AfqHam / AfqTrial pair."""

    def __init__(self, norb, nalpha, nbeta, ecore, h1, h1_mod, chol,
                 phi_up, phi_dn=None, trial_rhf=False, h1_dn=None, h1_mod_dn=None, chol_dn=None):
        """h1_dn, h1_mod_dn, chol_dn: unrestricted integrals (h1, h1_mod, chol are then
        the up-spin ones), or None."""
        self.norb, self.nalpha, self.nbeta = int(norb), int(nalpha), int(nbeta)
        self.ecore = float(ecore)
        self.h1 = np.ascontiguousarray(h1, dtype=np.float64).reshape(norb, norb)
        self.h1_mod = np.ascontiguousarray(h1_mod, dtype=np.float64).reshape(norb, norb)
        self.chol = np.ascontiguousarray(chol, dtype=np.float64).reshape(-1, norb, norb)
        self.nchol = self.chol.shape[0]
        self.unrestricted = chol_dn is not None
        if self.unrestricted:
            self.h1_dn = np.ascontiguousarray(h1_dn, dtype=np.float64).reshape(norb, norb)
            self.h1_mod_dn = np.ascontiguousarray(h1_mod_dn, dtype=np.float64).reshape(norb, norb)
            self.chol_dn = np.ascontiguousarray(chol_dn, dtype=np.float64).reshape(self.nchol, norb, norb)
        else:
            self.h1_dn = self.h1_mod_dn = self.chol_dn = None
        self.trial_rhf = bool(trial_rhf)
        self.phi_up = np.ascontiguousarray(phi_up, dtype=np.float64).reshape(norb, nalpha)
        if phi_dn is None:
            phi_dn = self.phi_up
        self.phi_dn = np.ascontiguousarray(phi_dn, dtype=np.float64).reshape(norb, -1)

    @property
    def mat_size(self):
        return self.norb * (self.nalpha + self.nbeta)

    def ham_struct(self):
        return AfqHam(self.norb, self.nalpha, self.nbeta, self.nchol, self.ecore,
                      dptr(self.h1), dptr(self.h1_mod), dptr(self.chol),
                      dptr(self.h1_dn), dptr(self.h1_mod_dn), dptr(self.chol_dn))

    def trial_struct(self):
        return AfqTrial(1 if self.trial_rhf else 0, dptr(self.phi_up), dptr(self.phi_dn))

    def serial_det(self, up, dn):
        """One walker in the serial layout (complex, length mat_size)."""
        up = np.asarray(up, dtype=np.complex128).reshape(self.norb, self.nalpha)
        dn = np.asarray(dn, dtype=np.complex128).reshape(self.norb, self.nbeta)
        return np.concatenate([up.ravel(), dn.ravel()])

    def split_det(self, serial):
        n, a, b = self.norb, self.nalpha, self.nbeta
        serial = np.asarray(serial)
        return serial[: n * a].reshape(n, a), serial[n * a: n * (a + b)].reshape(n, b)
