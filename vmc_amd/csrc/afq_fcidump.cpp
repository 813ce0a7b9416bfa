// FCIDUMP_chol (HDF5) readers and writers: readDQMCIntegralsRG / readDQMCIntegralsU
// (utils/integral.cpp:589-741) and write_dqmc / write_dqmc_uihf (scripts/QMCUtils.py:1067-1098),
// host-side over the HDF5 C library.
#include <hdf5.h>

#include <cstdint>
#include <vector>

#include "../../include/vmc_afqmc_io.h"

namespace {

struct File {
  hid_t id = -1;
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
};

struct Dataset {
  hid_t id = -1;
  ~Dataset() {
    if (id >= 0) H5Dclose(id);
  }
};

// number of elements of dataset `name`, or -1 when it is missing
int64_t dataset_size(hid_t file, const char* name) {
  Dataset d;
  H5E_BEGIN_TRY { d.id = H5Dopen2(file, name, H5P_DEFAULT); }
  H5E_END_TRY
  if (d.id < 0) return -1;
  const hid_t sp = H5Dget_space(d.id);
  if (sp < 0) return -1;
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  H5Sclose(sp);
  return (int64_t)n;
}

// reads dataset `name` (exactly n elements) as `type` into out
int read_dataset(hid_t file, const char* name, hid_t type, void* out, int64_t n) {
  const int64_t have = dataset_size(file, name);
  if (have < 0) return AFQ_IO_ERR_DATASET;
  if (have != n) return AFQ_IO_ERR_SHAPE;
  Dataset d;
  herr_t st = -1;
  H5E_BEGIN_TRY {
    d.id = H5Dopen2(file, name, H5P_DEFAULT);
    if (d.id >= 0) st = H5Dread(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
  }
  H5E_END_TRY
  return st < 0 ? AFQ_IO_ERR_DATASET : AFQ_IO_OK;
}

int open_ro(const char* path, File& f) {
  if (!path) return AFQ_IO_ERR_ARG;
  H5E_BEGIN_TRY { f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); }
  H5E_END_TRY
  return f.id < 0 ? AFQ_IO_ERR_OPEN : AFQ_IO_OK;
}

// a 1-D (n > 0) or scalar (n == 0) dataset of `type`
int write_dataset(hid_t file, const char* name, hid_t type, const void* data, int64_t n) {
  hid_t sp;
  if (n > 0) {
    const hsize_t dims[1] = {(hsize_t)n};
    sp = H5Screate_simple(1, dims, nullptr);
  } else {
    sp = H5Screate(H5S_SCALAR);
  }
  if (sp < 0) return AFQ_IO_ERR_DATASET;
  Dataset d;
  d.id = H5Dcreate2(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(sp);
  if (d.id < 0) return AFQ_IO_ERR_DATASET;
  return H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? AFQ_IO_ERR_DATASET : AFQ_IO_OK;
}

}  // namespace

extern "C" int afq_fcidump_header(const char* path, int64_t* header) {
  if (!header) return AFQ_IO_ERR_ARG;
  File f;
  int rc = open_ro(path, f);
  if (rc) return rc;
  return read_dataset(f.id, "/header", H5T_NATIVE_INT64, header, 4);
}

extern "C" int afq_fcidump_read(const char* path, double* h1, double* h1_mod, double* chol, double* ecore) {
  File f;
  int rc = open_ro(path, f);
  if (rc) return rc;
  int64_t hd[4];
  if ((rc = read_dataset(f.id, "/header", H5T_NATIVE_INT64, hd, 4))) return rc;
  const int64_t N = hd[1], G = hd[3];
  if (N <= 0 || G < 0) return AFQ_IO_ERR_SHAPE;
  if (h1 && (rc = read_dataset(f.id, "/hcore", H5T_NATIVE_DOUBLE, h1, N * N))) return rc;
  if (h1_mod && (rc = read_dataset(f.id, "/hcore_mod", H5T_NATIVE_DOUBLE, h1_mod, N * N))) return rc;
  if (chol && (rc = read_dataset(f.id, "/chol", H5T_NATIVE_DOUBLE, chol, G * N * N))) return rc;
  if (ecore) {
    *ecore = 0.0;  // the reference sets the core energy to zero when the dataset is absent
    if (dataset_size(f.id, "/energy_core") == 1) {
      double e = 0.0;
      if (read_dataset(f.id, "/energy_core", H5T_NATIVE_DOUBLE, &e, 1) == AFQ_IO_OK) *ecore = e;
    }
  }
  return AFQ_IO_OK;
}

extern "C" int afq_fcidump_read_u(const char* path, double* h1_up, double* h1_dn, double* h1_mod_up,
                                  double* h1_mod_dn, double* chol_up, double* chol_dn, double* ecore) {
  File f;
  int rc = open_ro(path, f);
  if (rc) return rc;
  int64_t hd[4];
  if ((rc = read_dataset(f.id, "/header", H5T_NATIVE_INT64, hd, 4))) return rc;
  const int64_t N = hd[1], G = hd[3];
  if (N <= 0 || G < 0) return AFQ_IO_ERR_SHAPE;
  struct Item { const char* name; double* out; int64_t n; };
  const Item items[6] = {{"/hcore_up", h1_up, N * N},         {"/hcore_dn", h1_dn, N * N},
                         {"/hcore_mod_up", h1_mod_up, N * N}, {"/hcore_mod_dn", h1_mod_dn, N * N},
                         {"/chol_up", chol_up, G * N * N},    {"/chol_dn", chol_dn, G * N * N}};
  for (const Item& it : items)
    if (it.out && (rc = read_dataset(f.id, it.name, H5T_NATIVE_DOUBLE, it.out, it.n))) return rc;
  if (ecore) {
    *ecore = 0.0;
    if (dataset_size(f.id, "/energy_core") == 1) {
      double e = 0.0;
      if (read_dataset(f.id, "/energy_core", H5T_NATIVE_DOUBLE, &e, 1) == AFQ_IO_OK) *ecore = e;
    }
  }
  return AFQ_IO_OK;
}

extern "C" int afq_fcidump_write(const char* path, int64_t nelec, int64_t norb, int64_t ms, int64_t nchol,
                                 const double* h1, const double* h1_mod, const double* chol, double ecore) {
  if (!path || !h1 || !h1_mod || (!chol && nchol > 0) || norb <= 0 || nchol < 0) return AFQ_IO_ERR_ARG;
  File f;
  f.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f.id < 0) return AFQ_IO_ERR_OPEN;
  const int64_t hd[4] = {nelec, norb, ms, nchol};
  const int64_t NN = norb * norb;
  int rc;
  if ((rc = write_dataset(f.id, "header", H5T_STD_I64LE, hd, 4))) return rc;
  if ((rc = write_dataset(f.id, "hcore", H5T_IEEE_F64LE, h1, NN))) return rc;
  if ((rc = write_dataset(f.id, "hcore_mod", H5T_IEEE_F64LE, h1_mod, NN))) return rc;
  if ((rc = write_dataset(f.id, "chol", H5T_IEEE_F64LE, chol, nchol * NN))) return rc;
  return write_dataset(f.id, "energy_core", H5T_IEEE_F64LE, &ecore, 0);
}

extern "C" int afq_fcidump_write_u(const char* path, int64_t nelec, int64_t norb, int64_t ms, int64_t nchol,
                                   const double* h1_up, const double* h1_dn, const double* h1_mod_up,
                                   const double* h1_mod_dn, const double* chol_up, const double* chol_dn,
                                   double ecore) {
  if (!path || !h1_up || !h1_dn || !h1_mod_up || !h1_mod_dn || norb <= 0 || nchol < 0 ||
      (nchol > 0 && (!chol_up || !chol_dn)))
    return AFQ_IO_ERR_ARG;
  File f;
  f.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f.id < 0) return AFQ_IO_ERR_OPEN;
  const int64_t hd[4] = {nelec, norb, ms, nchol};
  const int64_t NN = norb * norb;
  int rc;
  if ((rc = write_dataset(f.id, "header", H5T_STD_I64LE, hd, 4))) return rc;
  if ((rc = write_dataset(f.id, "hcore_up", H5T_IEEE_F64LE, h1_up, NN))) return rc;
  if ((rc = write_dataset(f.id, "hcore_dn", H5T_IEEE_F64LE, h1_dn, NN))) return rc;
  if ((rc = write_dataset(f.id, "hcore_mod_up", H5T_IEEE_F64LE, h1_mod_up, NN))) return rc;
  if ((rc = write_dataset(f.id, "hcore_mod_dn", H5T_IEEE_F64LE, h1_mod_dn, NN))) return rc;
  if ((rc = write_dataset(f.id, "chol_up", H5T_IEEE_F64LE, chol_up, nchol * NN))) return rc;
  if ((rc = write_dataset(f.id, "chol_dn", H5T_IEEE_F64LE, chol_dn, nchol * NN))) return rc;
  return write_dataset(f.id, "energy_core", H5T_IEEE_F64LE, &ecore, 0);
}
