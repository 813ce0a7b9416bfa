/*
 * vmc_afqmc_io.h -- the reference's Cholesky integral file (FCIDUMP_chol, HDF5)
 * read and written host-side (vmc_amd/libvmc_fcidump.so, linked against libhdf5).
 *
 * The file holds /header int64[4] = {nelec, norb, ms, nchol}, /hcore and
 * /hcore_mod f64[norb^2] (row-major), /chol f64[nchol norb^2] (row g = L_g,
 * symmetric) and /energy_core f64 scalar.  nalpha = (nelec + ms) / 2,
 * nbeta = nelec - nalpha.  The arrays feed afq_ham (include/vmc_afqmc.h)
 * unchanged.
 */
#ifndef VMC_AFQMC_IO_H
#define VMC_AFQMC_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFQ_IO_OK 0
#define AFQ_IO_ERR_OPEN (-1)     /* file missing or not HDF5 ("Cholesky integrals not found!") */
#define AFQ_IO_ERR_DATASET (-2)  /* a required dataset is missing or unreadable               */
#define AFQ_IO_ERR_SHAPE (-3)    /* a dataset's size does not match the header                */
#define AFQ_IO_ERR_ARG (-4)

/* readDQMCIntegralsRG's header read (utils/integral.cpp:604-616): header[4] =
 * {nelec, norb, ms, nchol}. */
int afq_fcidump_header(const char* path, int64_t* header);

/* readDQMCIntegralsRG (utils/integral.cpp:589-661): h1 and h1_mod norb x norb
 * row-major (readMat, utils/input.cpp:386-412), chol nchol x norb x norb, ecore
 * = /energy_core, or 0 when that dataset is absent (as the reference). */
int afq_fcidump_read(const char* path, double* h1, double* h1_mod, double* chol, double* ecore);

/* readDQMCIntegralsU (utils/integral.cpp:664-741), unrestricted integrals:
 * /hcore_up, /hcore_dn, /hcore_mod_up, /hcore_mod_dn (norb x norb row-major) and
 * /chol_up, /chol_dn (nchol x norb x norb).  Returns AFQ_IO_ERR_DATASET for a
 * restricted file. */
int afq_fcidump_read_u(const char* path, double* h1_up, double* h1_dn, double* h1_mod_up, double* h1_mod_dn,
                       double* chol_up, double* chol_dn, double* ecore);

/* write_dqmc (scripts/QMCUtils.py:1067-1075): the same datasets, types and
 * shapes the reference's writer produces. */
int afq_fcidump_write(const char* path, int64_t nelec, int64_t norb, int64_t ms, int64_t nchol,
                      const double* h1, const double* h1_mod, const double* chol, double ecore);

/* write_dqmc_uihf (scripts/QMCUtils.py:1088-1098): unrestricted integrals, the
 * datasets afq_fcidump_read_u reads. */
int afq_fcidump_write_u(const char* path, int64_t nelec, int64_t norb, int64_t ms, int64_t nchol,
                        const double* h1_up, const double* h1_dn, const double* h1_mod_up,
                        const double* h1_mod_dn, const double* chol_up, const double* chol_dn, double ecore);

#ifdef __cplusplus
}
#endif
#endif /* VMC_AFQMC_IO_H */
