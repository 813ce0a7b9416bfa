// calcMixedEstimatorLongProp (DQMC/MixedEstimator.cpp:309-699) over the C ABI:
// reference energy and Cholesky truncation, walker setup, the step loop with
// population control, periodic re-orthonormalisation, measurement with the
// E_L clip, and comb reconfiguration.  Single GPU; nranks virtual ranks model
// an `mpirun -np nranks` run of the reference (rank streams, rank-0 comb,
// rank-local running averages).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/vmc_afqmc.h"

void afq_keep_error(const afq_ctx* c, int rc);

namespace {

struct CtxGuard {
  afq_ctx* c = nullptr;
  int rc = AFQ_OK;
  ~CtxGuard() {
    if (c) {
      afq_keep_error(c, rc);
      afq_destroy(c);
    }
  }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" int afq_run_long_prop(const afq_ham* ham, const afq_trial* left, const double* ref_up,
                                 const double* ref_dn, const double* init_up, const double* init_dn,
                                 const afq_run_params* p, double* weights_out, double* energies_out,
                                 afq_run_result* res) {
  if (!ham || !left || !ref_up || !ref_dn || !init_up || !init_dn || !p) return AFQ_ERR_ARG;
  const int R = p->nranks > 0 ? p->nranks : 1, nw = p->nwalk, W = R * nw;
  const int N = ham->norb, na = ham->nalpha, nb = ham->nbeta, G = ham->nchol;
  const double dt = p->dt;
  afq_params cp{};
  cp.nwalk = W;
  cp.walker_rhf = p->walker_rhf;
  cp.ref_compat = p->ref_compat;
  cp.prec_fb = p->prec_fb;
  cp.prec_vhs = p->prec_vhs;
  cp.prec_exchange = p->prec_exchange;
  cp.device = p->device;
  // the reference's default cap uses the per-rank walker count (DQMC/MixedEstimator.cpp:452-454)
  cp.weight_cap = p->weight_cap > 0 ? p->weight_cap : std::max(100., nw / 10.);
  CtxGuard g;
  int& rc = g.rc;
  rc = afq_create(ham, left, &cp, &g.c);
  if (rc) return rc;
  afq_ctx* c = g.c;
  if (p->prec_prop != AFQ_PREC_FP64 && (rc = afq_set_prop_precision(c, p->prec_prop))) return rc;
  const size_t ms = (size_t)N * (na + nb);
  std::vector<double> ref(2 * ms, 0.0), init(2 * ms, 0.0);
  for (int i = 0; i < N; ++i) {
    for (int j = 0; j < na; ++j) {
      ref[2 * ((size_t)i * na + j)] = ref_up[(size_t)i * na + j];
      init[2 * ((size_t)i * na + j)] = init_up[(size_t)i * na + j];
    }
    for (int j = 0; j < nb; ++j) {
      const size_t q = (size_t)N * na + (size_t)i * nb + j;
      ref[2 * q] = ref_dn[(size_t)i * nb + j];
      init[2 * q] = init_dn[(size_t)i * nb + j];
    }
  }
  // reference energy and Cholesky truncation for E_L (:352-384)
  double o4[4];
  if ((rc = afq_ham_and_overlap(c, ref.data(), 1, G, o4))) return rc;
  auto cdivc = [](double ar, double ai, double br, double bi, double& cr, double& ci) {
    const double d = br * br + bi * bi;
    cr = (ar * br + ai * bi) / d;
    ci = (ai * br - ar * bi) / d;
  };
  double ref_re, ref_im;
  cdivc(o4[0], o4[1], o4[2], o4[3], ref_re, ref_im);
  const double ene0 = (p->ene0_guess == 1.e10) ? ref_re : p->ene0_guess;
  int nchol = G;
  double delta = 0.0;
  const int cand[5] = {int(0.3 * G), int(0.4 * G), int(0.5 * G), int(0.6 * G), int(0.7 * G)};
  for (int i = 0; i < 5; ++i) {
    if ((rc = afq_ham_and_overlap(c, ref.data(), 1, cand[i], o4))) return rc;
    double er, ei;
    cdivc(o4[0], o4[1], o4[2], o4[3], er, ei);
    if (std::hypot(ref_re - er, ref_im - ei) < p->cholesky_threshold) {
      nchol = cand[i];
      delta = ref_re - er;
      break;
    }
  }
  if ((rc = afq_set_nchol_ene(c, nchol))) return rc;
  const double* bg = p->rdm_background;
  if ((rc = afq_prep_prop(c, ref_up, ref_dn, bg, bg ? bg + (size_t)N * N : nullptr, dt, ene0))) return rc;
  std::vector<double> dets((size_t)W * 2 * ms);
  for (int w = 0; w < W; ++w) std::memcpy(&dets[(size_t)w * 2 * ms], init.data(), 2 * ms * sizeof(double));
  if ((rc = afq_set_walkers(c, dets.data(), nullptr, nullptr))) return rc;
  dets.clear();
  dets.shrink_to_fit();
  // Taylor tile shape for this population, timed once before the walk
  if ((rc = afq_autotune(c, nullptr, nullptr))) return rc;
  std::vector<double> eloc(W), weights(W, 1.0);
  const int nsweeps = p->nsweeps, nsteps = p->nsteps;
  std::vector<double> totalWeights(nsweeps, 0.0), totalEnergies(nsweeps, 0.0);
  if ((rc = afq_measure(c, AFQ_ENERGY_EXACT, 0.0, eloc.data(), nullptr))) return rc;
  double esum = 0.0;
  for (int w = 0; w < W; ++w) esum += eloc[w];
  totalWeights[0] = W;
  totalEnergies[0] = esum / W + delta;
  // random streams
  std::vector<std::mt19937> gens;
  for (int r = 0; r < R; ++r) gens.emplace_back((unsigned)(p->seed + r));
  std::vector<std::normal_distribution<double>> normals(W, std::normal_distribution<double>(0., 1.));
  std::vector<double> draws;
  if (p->rng_mode == AFQ_RNG_REPLAY) draws.assign((size_t)W * G, 0.0);
  double eshift = totalEnergies[0];
  double totalWeight = W;
  std::vector<double> averageEnergy(R, totalEnergies[0]), averageEnergyEql(R, totalEnergies[0]);
  double averageNum = 0., averageDenom = 0., averageNumEql = 0., averageDenomEql = 0.;
  double eEstimate = totalEnergies[0];
  int64_t nLarge = 0;
  double propTime = 0., eneTime = 0., tStart = 0.;
  int64_t timed = 0;
  afq_stage_times(c, nullptr, 1);
  if (p->one_rdm_out && (rc = afq_rdm_reset(c))) return rc;
  const bool cv = p->energy_mode == AFQ_ENERGY_SURROGATE;
  const bool mlp = cv && p->surrogate_kind == AFQ_SURR_MLP;
  const int ntrain = mlp ? std::max(1, p->surrogate_train_steps) : 0;
  double cv_var = 0.0, cv_rms = 0.0;
  int64_t cv_n = 0, nsamples = 0;
  double tr_rmse = 0.0, va_rmse = 0.0;
  if (cv) {
    if (mlp) {
      const int hidden = p->surrogate_hidden > 0 ? p->surrogate_hidden : 256;
      const int64_t cap = (int64_t)ntrain * W;
      if ((rc = afq_surrogate_init(c, hidden, (uint64_t)p->seed + 17, cap))) return rc;
    }
    if ((rc = afq_set_surrogate_kind(c, p->surrogate_kind, p->surrogate_prec))) return rc;
    if ((rc = afq_set_cv(c, p->subset_inv > 0 ? p->subset_inv : 16, p->risk_frac, p->fallback_rms,
                         (uint64_t)p->seed + 29)))
      return rc;
  }
  int64_t nmeasured = 0;
  double est_var = 0.0;
  int64_t est_n = 0;
  // shadow exact measurements of the surrogate blocks
  std::vector<double> eloc_sh(W), centers(W);
  double sh_sum = 0.0, sh_sum2 = 0.0, sh_est = 0.0, sh_wall = 0.0, sh_clip = 0.0;
  double sh_stage[AFQ_NSTAGES] = {0};
  int64_t sh_n = 0;
  double live_sum = 0.0;
  const double clip_w = std::sqrt(2. / dt);
  // DQMC/MixedEstimator.cpp:525-530: clip to the rank's running average -+ sqrt(2/dt)
  auto clip_host = [&](std::vector<double>& e, bool count) {
    for (int r = 0; r < R; ++r)
      for (int wl = 0; wl < nw; ++wl) {
        const int w = r * nw + wl;
        if (weights[w] == 0.) { e[w] = 0.; continue; }
        if (std::isnan(e[w]) || std::isinf(e[w])) return false;
        if (std::fabs(e[w] - averageEnergy[r]) > clip_w) {
          if (count) ++nLarge;
          e[w] = e[w] > averageEnergy[r] ? averageEnergy[r] + clip_w : averageEnergy[r] - clip_w;
        }
      }
    return true;
  };
  auto block_mean = [&](const std::vector<double>& e, double& var) {
    double we = 0.0, tw = 0.0;
    for (int w = 0; w < W; ++w) { we += e[w] * weights[w]; tw += weights[w]; }
    const double m = we / tw;
    double v = 0.0;
    for (int w = 0; w < W; ++w) v += weights[w] * weights[w] * (e[w] - m) * (e[w] - m);
    var = v / (tw * tw);
    return m;
  };
  const int64_t last = (int64_t)nsweeps * nsteps;
  for (int64_t step = 1; step < last; ++step) {
    if (p->max_steps > 0 && step > p->max_steps) break;
    if (step == p->warmup_steps + 1) {
      afq_sync(c);
      afq_stage_times(c, nullptr, 1);
      tStart = now_s();
    }
    for (int r = 0; r < R; ++r)
      if (step * dt < 10.) averageEnergy[r] = averageEnergyEql[r];
    double t0 = now_s();
    int64_t ncap = 0;
    if (p->rng_mode == AFQ_RNG_REPLAY) {
      for (int r = 0; r < R; ++r)
        for (int wl = 0; wl < nw; ++wl) {
          const int w = r * nw + wl;
          if (weights[w] > 1.e-8)
            for (int q = 0; q < G; ++q) draws[(size_t)w * G + q] = normals[w](gens[r]);
        }
      rc = afq_propagate(c, draws.data(), 0, (uint64_t)step, eshift, nullptr, &ncap);
    } else {
      rc = afq_propagate(c, nullptr, (uint64_t)p->seed, (uint64_t)step, eshift, nullptr, &ncap);
    }
    if (rc) return rc;
    nLarge += ncap;
    if ((rc = afq_get_weights(c, weights.data()))) return rc;
    propTime += now_s() - t0;
    totalWeight = 0.0;
    int nlive = 0;
    for (int w = 0; w < W; ++w) {
      totalWeight += weights[w];
      nlive += weights[w] > 0.0;
    }
    if (step > p->warmup_steps) live_sum += (double)nlive / W;
    eshift = eEstimate - 0.1 * std::log(totalWeight / (double)W) / dt;
    if (step % p->ortho_steps == 0)
      if ((rc = afq_orthogonalize(c))) return rc;
    if (step % nsteps == 0) {
      const int block = (int)(step / nsteps);
      t0 = now_s();
      // the 1-RDM of the walkers about to be measured (:492-510)
      if (p->one_rdm_out && step * dt > 10. && step > (int64_t)p->burn_iter * nsteps)
        if ((rc = afq_rdm_accumulate(c))) return rc;
      const bool exact_now = !cv || nmeasured < ntrain;
      // training pairs come from the FP64 path whatever the production precision
      if (cv && exact_now && (rc = afq_set_precision(c, AFQ_PREC_FP64, AFQ_PREC_FP64, AFQ_PREC_FP64))) return rc;
      if (!exact_now) {  // the surrogate estimate clips before its correction
        for (int w = 0; w < W; ++w) centers[w] = averageEnergy[w / nw];
        if ((rc = afq_set_energy_clip(c, centers.data(), clip_w))) return rc;
      }
      if ((rc = afq_measure(c, exact_now ? AFQ_ENERGY_EXACT : AFQ_ENERGY_SURROGATE, delta, eloc.data(), nullptr)))
        return rc;
      if (cv && exact_now && (rc = afq_set_precision(c, p->prec_fb, p->prec_vhs, p->prec_exchange))) return rc;
      bool clipped_on_device = false;
      if (!exact_now) {
        double st[6];
        afq_cv_stats(c, st);
        clipped_on_device = st[2] == 0.0;  // no fallback to the exact path
      }
      eneTime += now_s() - t0;
      ++nmeasured;
      if (!exact_now && p->shadow_exact) {
        // exact FP64 energies of the same walkers, outside the timings
        if ((rc = afq_sync(c))) return rc;
        const double ts = now_s();
        double before[AFQ_NSTAGES], after[AFQ_NSTAGES];
        afq_stage_times(c, before, 0);
        if ((rc = afq_set_precision(c, AFQ_PREC_FP64, AFQ_PREC_FP64, AFQ_PREC_FP64))) return rc;
        if ((rc = afq_measure(c, AFQ_ENERGY_EXACT, delta, eloc_sh.data(), nullptr))) return rc;
        if ((rc = afq_set_precision(c, p->prec_fb, p->prec_vhs, p->prec_exchange))) return rc;
        afq_stage_times(c, after, 0);
        if (step > p->warmup_steps)
          for (int i = 0; i < AFQ_NSTAGES; ++i) sh_stage[i] += after[i] - before[i];
        if (step > p->warmup_steps) {   // how much of the block the clip decides
          int nl = 0, nc = 0;
          for (int w = 0; w < W; ++w)
            if (weights[w] != 0.0) {
              ++nl;
              nc += std::fabs(eloc_sh[w] - averageEnergy[w / nw]) > clip_w;
            }
          sh_clip += nl ? (double)nc / nl : 0.0;
        }
        if (!clip_host(eloc_sh, false)) return rc = AFQ_ERR_NUMERIC;
        std::vector<double> ecv = eloc;
        if (!clipped_on_device && !clip_host(ecv, false)) return rc = AFQ_ERR_NUMERIC;
        double v_sh, v_cv;
        const double d = block_mean(ecv, v_cv) - block_mean(eloc_sh, v_sh);
        if (step > p->warmup_steps) {
          sh_sum += d;
          sh_sum2 += d * d;
          sh_est += v_sh;
          ++sh_n;
        }
        if (step > p->warmup_steps) sh_wall += now_s() - ts;
      }
      if (cv && exact_now) {
        if ((rc = afq_surrogate_collect(c, &nsamples))) return rc;
        if (nmeasured == ntrain) {
          const int epochs = p->surrogate_epochs > 0 ? p->surrogate_epochs : 20;
          if ((rc = afq_surrogate_train(c, epochs, 512, 1e-3, (uint64_t)p->seed + 31, &tr_rmse, &va_rmse))) return rc;
        }
      } else if (cv && step > p->warmup_steps) {
        double st[6];
        afq_cv_stats(c, st);
        cv_var += st[3];
        cv_rms += st[1];
        ++cv_n;
      }
      if (clipped_on_device) {
        for (int w = 0; w < W; ++w) {
          if (weights[w] == 0.) eloc[w] = 0.;
          else if (std::isnan(eloc[w]) || std::isinf(eloc[w])) return rc = AFQ_ERR_NUMERIC;
        }
      } else if (!clip_host(eloc, true)) {
        return rc = AFQ_ERR_NUMERIC;
      }
      double we = 0.0;
      for (int w = 0; w < W; ++w) we += eloc[w] * weights[w];
      totalWeights[block] = totalWeight;
      totalEnergies[block] = we / totalWeight;
      if (step > p->warmup_steps) {  // variance of this block's weighted mean over the population
        double v = 0.0;
        for (int w = 0; w < W; ++w) {
          const double d = eloc[w] - totalEnergies[block];
          v += weights[w] * weights[w] * d * d;
        }
        est_var += v / (totalWeight * totalWeight);
        ++est_n;
      }
      if (step * dt < 10. || step < (int64_t)p->burn_iter * nsteps) {
        averageNumEql += totalEnergies[block] * totalWeights[block];
        averageDenomEql += totalWeights[block];
        averageEnergyEql[0] = averageNumEql / averageDenomEql;
      } else {
        averageNum += totalEnergies[block] * totalWeights[block];
        averageDenom += totalWeights[block];
        averageEnergy[0] = averageNum / averageDenom;
      }
      eEstimate = 0.9 * eEstimate + 0.1 * totalEnergies[block];
      std::uniform_real_distribution<double> uniform(0., 1.);
      const double zeta = uniform(gens[0]);
      double tot = 0.0;
      if ((rc = afq_comb(c, zeta, &tot))) return rc;
      totalWeight = tot;
      for (int w = 0; w < W; ++w) weights[w] = tot / W;
    }
    if (step > p->warmup_steps) ++timed;
  }
  afq_sync(c);
  const double tEnd = now_s();
  double rdm_cw = 0.0;
  if (p->one_rdm_out && (rc = afq_get_one_rdm(c, left->rhf ? 1 : 0, p->one_rdm_out, &rdm_cw))) return rc;
  for (int i = 0; i < nsweeps; ++i) {
    if (weights_out) weights_out[i] = totalWeights[i];
    if (energies_out) energies_out[i] = totalEnergies[i];
  }
  if (res) {
    res->ref_energy = ref_re;
    res->nchol_ene = nchol;
    res->delta = delta;
    res->n_large_deviations = nLarge;
    res->prop_time = propTime;
    res->ene_time = eneTime;
    res->timed_steps = timed;
    res->timed_wall_s = timed > 0 ? tEnd - tStart - sh_wall : 0.0;
    afq_stage_times(c, res->stage_ms, 1);
    for (int i = 0; i < AFQ_NSTAGES; ++i) res->stage_ms[i] -= sh_stage[i];
    res->train_rmse = tr_rmse;
    res->valid_rmse = va_rmse;
    res->cv_var_sub = cv_n ? cv_var / cv_n : 0.0;
    res->cv_rms = cv_n ? cv_rms / cv_n : 0.0;
    double st[6] = {0, 0, 0, 0, 0, 0};
    if (cv) afq_cv_stats(c, st);
    res->cv_fallbacks = (int64_t)st[5];
    res->train_samples = nsamples;
    res->est_var = est_n ? est_var / est_n : 0.0;
    res->shadow_n = sh_n;
    res->shadow_shift = sh_n ? sh_sum / sh_n : 0.0;
    res->shadow_var = sh_n > 1 ? (sh_sum2 - sh_sum * sh_sum / sh_n) / (sh_n - 1) : 0.0;
    res->shadow_est_var = sh_n ? sh_est / sh_n : 0.0;
    res->shadow_wall_s = sh_wall;
    res->shadow_clip_frac = sh_n ? sh_clip / sh_n : 0.0;
    res->live_frac = timed ? live_sum / timed : 0.0;
    res->one_rdm_weight = rdm_cw;
  }
  return AFQ_OK;
}
