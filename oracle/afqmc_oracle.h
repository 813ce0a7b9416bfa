/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU FP64 restatement of the reference's phaseless AFQMC walker step
 * (sanshar/VMC: DQMC/DQMCWalker.cpp, DQMC/RHF.cpp, DQMC/UHF.cpp,
 * DQMC/Hamiltonian.cpp, DQMC/MixedEstimator.cpp:309-699), one walker at a time,
 * exactly in the reference's per-walker semantics.  Only tests/, the smoke
 * check of __graft_entry__ and bench.py's cpu_baseline leg may load it, and
 * only as the checker / the timed CPU baseline; the product library never
 * links it.
 *
 * Input structs are the ones of include/vmc_afqmc.h so that tests feed both
 * sides from the same buffers.
 */
#ifndef AFQMC_ORACLE_H
#define AFQMC_ORACLE_H

#include <stdint.h>
#include "../include/vmc_afqmc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const afq_ham* ham, const afq_trial* left, int walker_rhf,
                    int nwalk, int ref_compat);
void orc_destroy(orc_ctx* c);
void orc_set_nchol_ene(orc_ctx* c, int nchol_ene);
void orc_prep_prop(orc_ctx* c, const double* ref_up, const double* ref_dn,
                   const double* rdm_up, const double* rdm_dn, double dt, double ene0);
/* mf: nchol complex; expob: N x N row-major; prop_constant: complex */
void orc_get_prop_data(orc_ctx* c, double* mf, double* expob, double* prop_constant);

void orc_set_walkers(orc_ctx* c, const double* dets, const double* overlaps,
                     const double* weights);
void orc_get_walkers(orc_ctx* c, double* dets, double* overlaps, double* weights);

/* DQMCWalker::propagatePhaselessRG of walker w with the standard normals x
 * (nchol); returns the phaseless weight factor.  vhs_out (nullable) receives
 * the N x N complex propagator V (row-major) used in this step, fb_out
 * (nullable) the force bias (nchol complex). */
double orc_propagate_walker(orc_ctx* c, int w, const double* x, double eshift,
                            double* vhs_out, double* fb_out);
/* The population loop DQMC/MixedEstimator.cpp:461-468: live walkers only,
 * draws row w used for walker w, weight cap applied. */
void orc_propagate(orc_ctx* c, const double* draws, double eshift, double weight_cap,
                   double* factors, int64_t* n_capped);
void orc_orthogonalize(orc_ctx* c);
void orc_orthogonalize_walker(orc_ctx* c, int w);
void orc_ham_and_overlap_walker(orc_ctx* c, int w, double* out4);
void orc_force_bias_walker(orc_ctx* c, int w, double* fb);
void orc_overlap_walker(orc_ctx* c, int w, double* ovlp);
void orc_ham_and_overlap_dets(orc_ctx* c, const double* dets, int ndet,
                              int nchol_ene, double* out);
/* Green's function theta_s = psi_s (Phi_s^T psi_s)^-1 for slot s (N x n_s
 * complex row-major), and G_s = theta_s Phi_s^T (N x N complex row-major). */
void orc_green_walker(orc_ctx* c, int w, int slot, double* theta, double* green);
/* Exchange-only part of the local energy, 1/2 sum_s m_s sum_g tr(f f), and
 * the one-body part, for walker w (complex each). */
void orc_energy_parts_walker(orc_ctx* c, int w, double* e1, double* ej, double* ek);
void orc_comb(orc_ctx* c, double zeta, double* total_weight);
/* One measurement of the mixed-estimator 1-RDM (DQMC/MixedEstimator.cpp:492-510)
 * over the walkers of nonzero weight, in order, as the reference's running
 * average: rdm (row-major; spin_summed: one N x N block, RHF::oneRDM; else two,
 * up then down, UHF::oneRDM) and cw (cumulative weight) carry the state. */
void orc_one_rdm(orc_ctx* c, int spin_summed, double* rdm, double* cw);

/* Driver: calcMixedEstimatorLongProp with nranks virtual MPI ranks of nwalk
 * walkers each, rank r drawing from std::mt19937(seed + r). */
typedef struct orc_run_params {
  int64_t seed;
  double dt;
  int32_t nsteps;          /* steps per energy evaluation  */
  int32_t nsweeps;         /* stochasticIter               */
  int32_t ortho_steps;
  int32_t burn_iter;
  int32_t nwalk;           /* per rank                     */
  int32_t nranks;
  int32_t walker_rhf;
  int32_t ref_compat;
  double cholesky_threshold;
  double ene0_guess;       /* 1e10: use the reference energy */
  double weight_cap;       /* <= 0: default                  */
  int32_t max_steps;       /* > 0: stop after this many steps (partial run) */
  double* one_rdm_out;     /* non-null: the mixed-estimator 1-RDM (RHF trial: N x N
                            * spin-summed; UHF trial: up, down) */
} orc_run_params;

typedef struct orc_run_result {
  double ref_energy;
  int32_t nchol_ene;
  double delta;
  int64_t n_large_deviations;
  double prop_time, ene_time;
  double one_rdm_weight;   /* cumulativeWeight of the 1-RDM */
} orc_run_result;

/* ref_up/ref_dn: reference dets (rhf.txt or uhf.txt per walker type);
 * init_up/init_dn: right-trial sample every walker starts from.
 * weights_out/energies_out: nsweeps entries (samples.dat columns). */
int orc_run_long_prop(const afq_ham* ham, const afq_trial* left,
                      const double* ref_up, const double* ref_dn,
                      const double* init_up, const double* init_dn,
                      const orc_run_params* p, double* weights_out,
                      double* energies_out, orc_run_result* res);

/* Reference blocking analysis (DQMC/MixedEstimator.cpp:15-49): out = {mean, error}. */
void orc_blocking(const double* weights, const double* energies, int n, double* out2);

/* CPU baseline: nthreads workers, each stepping its own shard of walkers with
 * its own std::mt19937(seed + worker) stream, for nsteps steps with E_L every
 * energy_every steps (0: never) and QR every ortho_steps.  Returns walker-steps
 * per second over the whole pool. */
double orc_bench(orc_ctx* c, int nthreads, int nsteps, int energy_every,
                 int ortho_steps, uint64_t seed, double* seconds);

/* libstdc++ stream helpers for tests: n draws of std::normal_distribution<double>
 * from a fresh std::mt19937(seed); and the 10000th output of std::mt19937(). */
void orc_normals(uint64_t seed, int n, double* out);
uint32_t orc_mt19937_10000th(void);

#ifdef __cplusplus
}
#endif
#endif
