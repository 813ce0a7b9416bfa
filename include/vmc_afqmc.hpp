/*
 * vmc_afqmc.hpp -- C++ mirror of the reference's walker-step interfaces over the
 * C ABI of vmc_afqmc.h (header only, C++17, no dependencies beyond the standard
 * library).
 *
 * The reference drives one DQMCWalker object per walker through a Wavefunction
 * (RHF / UHF) and a Hamiltonian (DQMC/DQMCWalker.h:7-46, DQMC/Wavefunction.h:7-20,
 * DQMC/Hamiltonian.h:8-40).  Here one DQMCWalkers object is the whole population of
 * a rank on one GPU; its methods keep the reference's names and argument meaning
 * and act on every walker at once:
 *
 *   reference (per walker)                       this header (population)
 *   Hamiltonian(fname) + setNcholEne             Hamiltonian(arrays...) + setNcholEne
 *   RHF(ham, leftQ) / UHF(ham, leftQ)            RHF(ham, det) / UHF(ham, up, dn)
 *   DQMCWalker::prepProp(ref, ham, dt, ene0)     DQMCWalkers::prepProp(ref_up, ref_dn, dt, ene0)
 *   DQMCWalker::setDet(serial, overlap)          DQMCWalkers::setDet(serial, overlaps, weights)
 *   DQMCWalker::getDet(serial) -> overlap        DQMCWalkers::getDet(serial, overlaps, weights)
 *   DQMCWalker::propagatePhaseless(wave, ham, e) DQMCWalkers::propagatePhaseless(draws, eshift)
 *   DQMCWalker::orthogonalize()                  DQMCWalkers::orthogonalize()
 *   DQMCWalker::overlap(wave)                    DQMCWalkers::overlap()
 *   DQMCWalker::forceBias(wave, ham, fb)         DQMCWalkers::forceBias()
 *   DQMCWalker::hamAndOverlap(wave, ham)         DQMCWalkers::hamAndOverlap()
 *   reconfiguration (MixedEstimator.cpp:559-635) DQMCWalkers::comb(zeta)
 *   calcMixedEstimatorLongProp                   calcMixedEstimatorLongProp(...)
 *
 * Errors throw vmc_afqmc::Error (the reference calls exit()).
 */
#ifndef VMC_AFQMC_HPP
#define VMC_AFQMC_HPP

#include <array>
#include <complex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vmc_afqmc.h"

namespace vmc_afqmc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const afq_ctx* ctx, const char* call) {
  if (rc != AFQ_OK) throw Error(rc, std::string(call) + ": " + afq_last_error(ctx));
}

using cplx = std::complex<double>;

// Cholesky Hamiltonian in the fields of FCIDUMP_chol (utils/integral.cpp:589-661).
// Matrices are row-major; chol holds G rows vec(L_g), each L_g symmetric N x N.
class Hamiltonian {
 public:
  int norbs, nalpha, nbeta, nchol, ncholEne;
  double ecore;
  std::vector<double> h1, h1Mod, chol;

  Hamiltonian(int norbs_, int nalpha_, int nbeta_, int nchol_, double ecore_, std::vector<double> h1_,
              std::vector<double> h1Mod_, std::vector<double> chol_)
      : norbs(norbs_), nalpha(nalpha_), nbeta(nbeta_), nchol(nchol_), ncholEne(nchol_), ecore(ecore_),
        h1(std::move(h1_)), h1Mod(std::move(h1Mod_)), chol(std::move(chol_)) {
    const size_t NN = (size_t)norbs * norbs;
    if (h1.size() != NN || h1Mod.size() != NN || chol.size() != NN * nchol)
      throw Error(AFQ_ERR_ARG, "Hamiltonian: array sizes do not match norbs / nchol");
  }
  // DQMC/Hamiltonian.cpp:45-48
  void setNcholEne(int n) { ncholEne = n; }

  afq_ham c_struct() const {
    afq_ham h{};  // restricted integrals: the down-spin pointers stay null
    h.norb = norbs; h.nalpha = nalpha; h.nbeta = nbeta; h.nchol = nchol; h.ecore = ecore;
    h.h1 = h1.data(); h.h1_mod = h1Mod.data(); h.chol = chol.data();
    return h;
  }
};

// Left trial: real orbitals, row-major N x n.
class Wavefunction {
 public:
  bool rhf;
  std::vector<double> up, dn;
  afq_trial c_struct() const {
    afq_trial t;
    t.rhf = rhf ? 1 : 0;
    t.phi_up = up.data();
    t.phi_dn = rhf ? up.data() : dn.data();
    return t;
  }

 protected:
  Wavefunction(bool r, std::vector<double> u, std::vector<double> d) : rhf(r), up(std::move(u)), dn(std::move(d)) {}
};

// DQMC/RHF.h: one set of orbitals for both spins (nalpha == nbeta).
class RHF : public Wavefunction {
 public:
  RHF(const Hamiltonian& ham, std::vector<double> det) : Wavefunction(true, std::move(det), {}) {
    if (ham.nalpha != ham.nbeta || up.size() != (size_t)ham.norbs * ham.nalpha)
      throw Error(AFQ_ERR_ARG, "RHF: needs nalpha == nbeta and an N x nalpha determinant");
  }
};

// DQMC/UHF.h
class UHF : public Wavefunction {
 public:
  UHF(const Hamiltonian& ham, std::vector<double> detUp, std::vector<double> detDn)
      : Wavefunction(false, std::move(detUp), std::move(detDn)) {
    if (up.size() != (size_t)ham.norbs * ham.nalpha || dn.size() != (size_t)ham.norbs * ham.nbeta)
      throw Error(AFQ_ERR_ARG, "UHF: determinant sizes do not match the Hamiltonian");
  }
};

// The walker population of one rank on one GPU.
class DQMCWalkers {
 public:
  // rhfQ: the walkers carry one determinant for both spins (DQMCWalker::rhfQ).
  DQMCWalkers(const Hamiltonian& ham, const Wavefunction& left, int nwalk, bool rhfQ, int device = 0,
              double weightCap = 0.0)
      : nwalk_(nwalk), matSize_((size_t)ham.norbs * (ham.nalpha + ham.nbeta)), nchol_(ham.nchol),
        norb_((size_t)ham.norbs) {
    afq_ham h = ham.c_struct();
    afq_trial t = left.c_struct();
    afq_params p{};
    p.nwalk = nwalk;
    p.walker_rhf = rhfQ ? 1 : 0;
    p.ref_compat = 0;
    p.prec_fb = p.prec_vhs = p.prec_exchange = AFQ_PREC_FP64;
    p.device = device;
    p.weight_cap = weightCap;
    check(afq_create(&h, &t, &p, &ctx_), nullptr, "afq_create");
    if (ham.ncholEne != ham.nchol) check(afq_set_nchol_ene(ctx_, ham.ncholEne), ctx_, "afq_set_nchol_ene");
  }
  DQMCWalkers(const DQMCWalkers&) = delete;
  DQMCWalkers& operator=(const DQMCWalkers&) = delete;
  ~DQMCWalkers() { afq_destroy(ctx_); }

  int size() const { return nwalk_; }
  size_t matSize() const { return matSize_; }
  afq_ctx* handle() { return ctx_; }

  // DQMC/DQMCWalker.cpp:32-103 (reference determinant -> background, exp(-dt/2 h1'))
  void prepProp(const std::vector<double>& refUp, const std::vector<double>& refDn, double dt, double ene0) {
    check(afq_prep_prop(ctx_, refUp.data(), refDn.data(), nullptr, nullptr, dt, ene0), ctx_, "afq_prep_prop");
  }
  // DQMC/DQMCWalker.cpp:251-269, for every walker; overlaps may be empty (computed)
  void setDet(const std::vector<cplx>& serial, const std::vector<cplx>& overlaps = {},
              const std::vector<double>& weights = {}) {
    if (serial.size() != matSize_ * nwalk_) throw Error(AFQ_ERR_ARG, "setDet: serial size");
    check(afq_set_walkers(ctx_, reinterpret_cast<const double*>(serial.data()),
                          overlaps.empty() ? nullptr : reinterpret_cast<const double*>(overlaps.data()),
                          weights.empty() ? nullptr : weights.data()),
          ctx_, "afq_set_walkers");
  }
  // DQMC/DQMCWalker.cpp:272-290
  void getDet(std::vector<cplx>& serial, std::vector<cplx>& overlaps, std::vector<double>& weights) {
    serial.resize(matSize_ * nwalk_);
    overlaps.resize(nwalk_);
    weights.resize(nwalk_);
    check(afq_get_walkers(ctx_, reinterpret_cast<double*>(serial.data()), reinterpret_cast<double*>(overlaps.data()),
                          weights.data()),
          ctx_, "afq_get_walkers");
  }
  std::vector<double> weights() {
    std::vector<double> w(nwalk_);
    check(afq_get_weights(ctx_, w.data()), ctx_, "afq_get_weights");
    return w;
  }
  // DQMC/DQMCWalker.cpp:441-503 and the weight update of DQMC/MixedEstimator.cpp:461-468 for
  // every live walker.  draws: nwalk x nchol normals (the reference's stream), or empty for the
  // device generator (seed, step).  Returns the weight factors; nCapped receives the walkers
  // zeroed by the weight cap.
  std::vector<double> propagatePhaseless(const std::vector<double>& draws, double eshift, uint64_t seed = 0,
                                         uint64_t step = 0, int64_t* nCapped = nullptr) {
    if (!draws.empty() && draws.size() != (size_t)nwalk_ * nchol_) throw Error(AFQ_ERR_ARG, "draws size");
    std::vector<double> f(nwalk_);
    int64_t nc = 0;
    check(afq_propagate(ctx_, draws.empty() ? nullptr : draws.data(), seed, step, eshift, f.data(), &nc), ctx_,
          "afq_propagate");
    if (nCapped) *nCapped = nc;
    return f;
  }
  // DQMC/DQMCWalker.cpp:295-326
  void orthogonalize() { check(afq_orthogonalize(ctx_), ctx_, "afq_orthogonalize"); }
  // DQMC/DQMCWalker.cpp:559-566
  std::vector<cplx> overlap() {
    std::vector<cplx> o(nwalk_);
    check(afq_overlap(ctx_, reinterpret_cast<double*>(o.data())), ctx_, "afq_overlap");
    return o;
  }
  // DQMC/DQMCWalker.cpp:569-576: nwalk x nchol
  std::vector<cplx> forceBias() {
    std::vector<cplx> fb((size_t)nwalk_ * nchol_);
    check(afq_force_bias(ctx_, reinterpret_cast<double*>(fb.data())), ctx_, "afq_force_bias");
    return fb;
  }
  // DQMC/DQMCWalker.cpp:592-601: {E*O, O} of every walker (dead walkers: E*O = 0)
  std::vector<std::array<cplx, 2>> hamAndOverlap() {
    std::vector<std::array<cplx, 2>> ho(nwalk_);
    check(afq_measure(ctx_, AFQ_ENERGY_EXACT, 0.0, nullptr, reinterpret_cast<double*>(ho.data())), ctx_,
          "afq_measure");
    return ho;
  }
  // local energies Re(E*O/O) + delta of the live walkers (0 for dead ones)
  std::vector<double> localEnergy(double delta = 0.0) {
    std::vector<double> e(nwalk_);
    check(afq_measure(ctx_, AFQ_ENERGY_EXACT, delta, e.data(), nullptr), ctx_, "afq_measure");
    return e;
  }
  // DQMC/MixedEstimator.cpp:559-635; returns the total weight before the comb
  double comb(double zeta) {
    double tot = 0.0;
    check(afq_comb(ctx_, zeta, &tot), ctx_, "afq_comb");
    return tot;
  }
  // DQMC/MixedEstimator.cpp:492-510 (DQMCWalker::oneRDM of every live walker) for one
  // measurement; oneRDM() returns the weighted average, N x N row-major per block
  // (spinSummed: one block, RHF::oneRDM; else up then down, UHF::oneRDM)
  void accumulateOneRDM() { check(afq_rdm_accumulate(ctx_), ctx_, "afq_rdm_accumulate"); }
  std::vector<double> oneRDM(bool spinSummed, double* cumulativeWeight = nullptr) {
    std::vector<double> r((spinSummed ? 1 : 2) * norb_ * norb_);
    check(afq_get_one_rdm(ctx_, spinSummed ? 1 : 0, r.data(), cumulativeWeight), ctx_, "afq_get_one_rdm");
    return r;
  }

 private:
  afq_ctx* ctx_ = nullptr;
  int nwalk_;
  size_t matSize_;
  int nchol_;
  size_t norb_;
};

// DQMC/MixedEstimator.cpp:309-699 on one GPU.  Returns the per-block
// (total weight, energy) pairs of samples.dat; result may be null.
inline std::vector<std::pair<double, double>> calcMixedEstimatorLongProp(
    const Hamiltonian& ham, const Wavefunction& left, const std::vector<double>& refUp,
    const std::vector<double>& refDn, const std::vector<double>& initUp, const std::vector<double>& initDn,
    const afq_run_params& params, afq_run_result* result = nullptr) {
  afq_ham h = ham.c_struct();
  afq_trial t = left.c_struct();
  std::vector<double> w(params.nsweeps), e(params.nsweeps);
  afq_run_result r{};
  check(afq_run_long_prop(&h, &t, refUp.data(), refDn.data(), initUp.data(), initDn.data(), &params, w.data(),
                          e.data(), &r),
        nullptr, "afq_run_long_prop");
  if (result) *result = r;
  std::vector<std::pair<double, double>> out(params.nsweeps);
  for (int i = 0; i < params.nsweeps; ++i) out[i] = {w[i], e[i]};
  return out;
}

}  // namespace vmc_afqmc
#endif
