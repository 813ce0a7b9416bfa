// ORACLE -- test infrastructure, not product code.  See afqmc_oracle.h.
//
// CPU FP64 restatement of the reference's phaseless AFQMC walker step, written
// from the equations of the reference's source (sanshar/VMC), one walker at a
// time.  Dense linear algebra is this file's own (column-major like Eigen):
// partial-pivoting LU (determinant, inverse), Householder QR with Eigen's sign
// convention, and the symmetric matrix exponential by Jacobi diagonalisation.
//
// ref_compat reproduces the reference's FP32 pieces: L rounded to float in the
// mean-field shifts / one-body operator (DQMC/DQMCWalker.cpp:73-84, because the
// left trial sets Hamiltonian::rotFlag, DQMC/Hamiltonian.cpp:83,132), fields
// rounded to complex<float> (DQMC/DQMCWalker.cpp:454) and the VHS result
// rounded to float (DQMC/DQMCWalker.cpp:406-421; the sum itself is formed in
// double here, so it is not Eigen's float summation order).
#include "afqmc_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

namespace {

using cd = std::complex<double>;

template <class T>
struct Mat {
  int r = 0, c = 0;
  std::vector<T> a;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, T(0)) {}
  T& operator()(int i, int j) { return a[(size_t)i + (size_t)j * r]; }
  const T& operator()(int i, int j) const { return a[(size_t)i + (size_t)j * r]; }
};
using MC = Mat<cd>;
using MD = Mat<double>;

// C = A * B, A real (m x k), B complex (k x n)
MC mul(const MD& A, const MC& B) {
  MC C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int k = 0; k < A.c; ++k) {
      const cd b = B(k, j);
      const double* a = &A.a[(size_t)k * A.r];
      cd* c = &C.a[(size_t)j * C.r];
      for (int i = 0; i < A.r; ++i) c[i] += a[i] * b;
    }
  return C;
}

// C = A * B, complex
MC mul(const MC& A, const MC& B) {
  MC C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int k = 0; k < A.c; ++k) {
      const cd b = B(k, j);
      const cd* a = &A.a[(size_t)k * A.r];
      cd* c = &C.a[(size_t)j * C.r];
      for (int i = 0; i < A.r; ++i) c[i] += a[i] * b;
    }
  return C;
}

MD transpose(const MD& A) {
  MD T(A.c, A.r);
  for (int j = 0; j < A.c; ++j)
    for (int i = 0; i < A.r; ++i) T(j, i) = A(i, j);
  return T;
}

// Partial-pivoting LU in place (the decomposition behind Eigen's
// determinant() / inverse() for dynamic sizes, eigen/Eigen/src/LU/Determinant.h:28-35,
// InverseImpl.h:23-29).  Pivot = largest modulus in the column.
void lu_factor(MC& A, std::vector<int>& piv, int& nswap) {
  const int n = A.r;
  piv.assign(n, 0);
  nswap = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(A(k, k));
    for (int i = k + 1; i < n; ++i) {
      double v = std::abs(A(i, k));
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (p != k) {
      ++nswap;
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
    }
    const cd d = A(k, k);
    if (d != cd(0.0)) {
      for (int i = k + 1; i < n; ++i) A(i, k) /= d;
    }
    for (int j = k + 1; j < n; ++j) {
      const cd akj = A(k, j);
      for (int i = k + 1; i < n; ++i) A(i, j) -= A(i, k) * akj;
    }
  }
}

cd lu_det(const MC& LU, int nswap) {
  cd d = (nswap % 2) ? cd(-1.0) : cd(1.0);
  for (int k = 0; k < LU.r; ++k) d *= LU(k, k);
  return d;
}

MC lu_inverse(const MC& LU, const std::vector<int>& piv) {
  const int n = LU.r;
  // permutation applied to identity rows in pivot order
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) std::swap(perm[k], perm[piv[k]]);
  MC X(n, n);
  for (int j = 0; j < n; ++j) {
    // b = P e_j : b_i = 1 if perm[i] == j
    std::vector<cd> b(n, cd(0.0));
    for (int i = 0; i < n; ++i)
      if (perm[i] == j) b[i] = 1.0;
    for (int i = 0; i < n; ++i)  // forward, unit lower
      for (int k = 0; k < i; ++k) b[i] -= LU(i, k) * b[k];
    for (int i = n - 1; i >= 0; --i) {  // backward, upper
      for (int k = i + 1; k < n; ++k) b[i] -= LU(i, k) * b[k];
      b[i] /= LU(i, i);
    }
    for (int i = 0; i < n; ++i) X(i, j) = b[i];
  }
  return X;
}

// Thin Householder QR with Eigen's conventions (makeHouseholder:
// eigen/Eigen/src/Householder/Householder.h:63-92, householder_qr_inplace_unblocked:
// eigen/Eigen/src/QR/HouseholderQR.h:267-297, householderQ() with conjugated
// coefficients :163-167).  A (m x n, m >= n) is replaced by the thin Q; returns
// prod_k R_kk (each R_kk = beta_k real).
cd householder_thin_q(MC& A) {
  const int m = A.r, n = A.c;
  std::vector<cd> tau(n);
  cd prod(1.0, 0.0);
  const double tol = std::numeric_limits<double>::min();
  for (int k = 0; k < n; ++k) {
    double tail = 0.0;
    for (int i = k + 1; i < m; ++i) tail += std::norm(A(i, k));
    const cd c0 = A(k, k);
    double beta;
    if (tail <= tol && std::imag(c0) * std::imag(c0) <= tol) {
      tau[k] = 0.0;
      beta = std::real(c0);
      for (int i = k + 1; i < m; ++i) A(i, k) = 0.0;
    } else {
      beta = std::sqrt(std::norm(c0) + tail);
      if (std::real(c0) >= 0.0) beta = -beta;
      const cd den = c0 - beta;
      for (int i = k + 1; i < m; ++i) A(i, k) /= den;
      tau[k] = std::conj((beta - c0) / beta);
    }
    A(k, k) = beta;
    prod *= beta;
    // apply H = I - tau v v^* from the left to the trailing columns
    for (int j = k + 1; j < n; ++j) {
      cd t = A(k, j);
      for (int i = k + 1; i < m; ++i) t += std::conj(A(i, k)) * A(i, j);
      A(k, j) -= tau[k] * t;
      for (int i = k + 1; i < m; ++i) A(i, j) -= tau[k] * A(i, k) * t;
    }
  }
  // Q = H_0^* H_1^* ... H_{n-1}^* [I; 0],  H_k^* = I - conj(tau_k) v_k v_k^*
  MC Q(m, n);
  for (int j = 0; j < n; ++j) Q(j, j) = 1.0;
  for (int k = n - 1; k >= 0; --k) {
    const cd tk = std::conj(tau[k]);
    if (tk == cd(0.0)) continue;
    for (int j = k; j < n; ++j) {
      cd t = Q(k, j);
      for (int i = k + 1; i < m; ++i) t += std::conj(A(i, k)) * Q(i, j);
      Q(k, j) -= tk * t;
      for (int i = k + 1; i < m; ++i) Q(i, j) -= tk * A(i, k) * t;
    }
  }
  A = Q;
  return prod;
}

// Cyclic Jacobi eigen-decomposition of a real symmetric matrix: A = V diag(w) V^T.
void jacobi_eig(MD A, MD& V, std::vector<double>& w) {
  const int n = A.r;
  V = MD(n, n);
  for (int i = 0; i < n; ++i) V(i, i) = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += A(i, j) * A(i, j);
        if (i != j) off += A(i, j) * A(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (std::abs(apq) < 1e-300) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::abs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V(k, p), vkq = V(k, q);
          V(k, p) = c * vkp - s * vkq;
          V(k, q) = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A(i, i);
}

// exp(s * H) for real symmetric H (the reference uses Eigen's Pade expm,
// DQMC/DQMCWalker.cpp:102; for a symmetric argument the closed form is exact).
MD expm_sym(const MD& H, double s) {
  MD V;
  std::vector<double> w;
  jacobi_eig(H, V, w);
  const int n = H.r;
  MD E(n, n);
  for (int k = 0; k < n; ++k) {
    const double e = std::exp(s * w[k]);
    for (int j = 0; j < n; ++j) {
      const double vjk = V(j, k) * e;
      for (int i = 0; i < n; ++i) E(i, j) += V(i, k) * vjk;
    }
  }
  return E;
}

struct Slot {
  int n;     // electrons in this spin slot
  int det;   // walker determinant used
  int phi;   // trial orbital set used
  int mult;  // multiplicity (2 when an RHF walker meets an RHF trial)
};

struct Walker {
  MC det[2];
  cd trialOverlap{1.0, 0.0};
  cd orthoFac{1.0, 0.0};
  std::normal_distribution<double> normal{0.0, 1.0};
};

}  // namespace

struct orc_ctx {
  int N = 0, na = 0, nb = 0, G = 0, GE = 0;
  double ecore = 0.0;
  MD h1, h1mod;
  std::vector<double> L;   // G x N x N (symmetric L_g); the up-spin vectors with unrestricted integrals
  std::vector<float> Lf;   // floattenCholesky: g*T + i(i+1)/2 + j, j <= i
  // unrestricted integrals (intType "u", utils/integral.cpp:664-741): the down-spin
  // one-body matrices, Cholesky vectors (and their float copy) and propagator
  bool uints = false;
  MD h1d, h1modd, expOBd;
  std::vector<double> Ld;
  std::vector<float> Lfd;
  bool trial_rhf = true, walker_rhf = true, ref_compat = false;
  MD phi[2], phiT[2];
  std::vector<double> R[2];  // rotCholMat[s]: R[s][g*(n*N) + k*n + i] = (Phi^T L_g)(i,k)
  std::vector<Slot> slots;
  double dt = 0.0;
  std::vector<cd> mf;
  cd propConstant{0.0, 0.0};
  MD expOB;
  bool prepped = false;
  std::vector<Walker> walkers;
  std::vector<double> weights;

  double Lval(int g, int i, int j) const {
    return L[(size_t)g * N * N + (size_t)i * N + j];
  }
  double Lfval(int g, int i, int j) const {
    if (j > i) std::swap(i, j);
    const size_t T = (size_t)N * (N + 1) / 2;
    return (double)Lf[(size_t)g * T + (size_t)i * (i + 1) / 2 + j];
  }
  // spin s's vectors: the shared ones with restricted integrals
  double Lval_s(int s, int g, int i, int j) const {
    return (uints && s == 1) ? Ld[(size_t)g * N * N + (size_t)i * N + j] : Lval(g, i, j);
  }
  double Lfval_s(int s, int g, int i, int j) const {
    if (!(uints && s == 1)) return Lfval(g, i, j);
    if (j > i) std::swap(i, j);
    const size_t T = (size_t)N * (N + 1) / 2;
    return (double)Lfd[(size_t)g * T + (size_t)i * (i + 1) / 2 + j];
  }
};

namespace {

void build_slots(orc_ctx* c) {
  c->slots.clear();
  if (c->walker_rhf && c->trial_rhf) {
    c->slots.push_back({c->na, 0, 0, 2});
  } else {
    c->slots.push_back({c->na, 0, 0, 1});
    c->slots.push_back({c->nb, c->walker_rhf ? 0 : 1, c->trial_rhf ? 0 : 1, 1});
  }
}

// theta_s = psi (Phi_s^T psi)^-1, det_s (DQMC/RHF.cpp:27-61, DQMC/UHF.cpp:31-53)
void green_slot(const orc_ctx* c, const Walker& W, const Slot& s, MC* theta, cd* det) {
  MC S = mul(c->phiT[s.phi], W.det[s.det]);
  std::vector<int> piv;
  int nswap;
  lu_factor(S, piv, nswap);
  if (det) *det = lu_det(S, nswap);
  if (theta) *theta = mul(W.det[s.det], lu_inverse(S, piv));
}

cd overlap(const orc_ctx* c, const Walker& W) {
  cd o(1.0, 0.0);
  for (const Slot& s : c->slots) {
    cd d;
    green_slot(c, W, s, nullptr, &d);
    o *= (s.mult == 2) ? d * d : d;
  }
  return o;
}

// fb_g = sum_s m_s sum_{k,i} theta_s(k,i) R_{s,g}(i,k)   (thetaTFlat^T rotCholMat)
void force_bias(const orc_ctx* c, const Walker& W, std::vector<cd>& fb) {
  fb.assign(c->G, cd(0.0));
  for (const Slot& s : c->slots) {
    MC th;
    green_slot(c, W, s, &th, nullptr);
    const int n = s.n, N = c->N;
    std::vector<cd> tflat((size_t)n * N);  // idx = k*n + i
    for (int k = 0; k < N; ++k)
      for (int i = 0; i < n; ++i) tflat[(size_t)k * n + i] = th(k, i);
    const size_t K = (size_t)n * N;
    const double* R = c->R[s.phi].data();
    for (int g = 0; g < c->G; ++g) {
      const double* r = R + (size_t)g * K;
      double re = 0, im = 0;
      for (size_t q = 0; q < K; ++q) {
        re += tflat[q].real() * r[q];
        im += tflat[q].imag() * r[q];
      }
      fb[g] += (double)s.mult * cd(re, im);
    }
  }
}

// {E*O, O} of DQMC/RHF.cpp:104-131 / DQMC/UHF.cpp:75-106 for ncholEne = GE.
void ham_and_overlap(const orc_ctx* c, const Walker& W, int GE, cd out[2],
                     cd* e1_out = nullptr, cd* ej_out = nullptr, cd* ek_out = nullptr) {
  const int N = c->N;
  cd O(1.0, 0.0);
  cd e1(0.0, 0.0);
  std::vector<MC> theta(c->slots.size());
  for (size_t si = 0; si < c->slots.size(); ++si) {
    const Slot& s = c->slots[si];
    cd d;
    green_slot(c, W, s, &theta[si], &d);
    O *= (s.mult == 2) ? d * d : d;
    // sum_ij G(i,j) h1(i,j), G = theta Phi^T
    const MD& P = c->phi[s.phi];
    // unrestricted integrals: spin s's h1 (DQMC/UHF.cpp: the intType "u" one-body term)
    const MD& h1 = (c->uints && si == 1) ? c->h1d : c->h1;
    cd acc(0.0, 0.0);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        cd g(0.0, 0.0);
        for (int k = 0; k < s.n; ++k) g += theta[si](i, k) * P(j, k);
        acc += g * h1(i, j);
      }
    e1 += (double)s.mult * acc;
  }
  cd ej(0.0, 0.0), ek(0.0, 0.0);
  std::vector<cd> f;
  for (int g = 0; g < GE; ++g) {
    cd csum(0.0, 0.0), X(0.0, 0.0);
    for (size_t si = 0; si < c->slots.size(); ++si) {
      const Slot& s = c->slots[si];
      const int n = s.n;
      const double* R = c->R[s.phi].data() + (size_t)g * n * N;  // col-major n x N
      f.assign((size_t)n * n, cd(0.0));                            // col-major n x n
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < N; ++k) {
          const cd t = theta[si](k, j);
          const double* r = R + (size_t)k * n;
          cd* fj = &f[(size_t)j * n];
          for (int i = 0; i < n; ++i) fj[i] += r[i] * t;
        }
      cd tr(0.0, 0.0), x(0.0, 0.0);
      for (int i = 0; i < n; ++i) {
        tr += f[(size_t)i * n + i];
        for (int j = 0; j < n; ++j) x += f[(size_t)j * n + i] * f[(size_t)i * n + j];
      }
      csum += (double)s.mult * tr;
      X += (double)s.mult * x;
    }
    ej += csum * csum / 2.0;
    ek += X / 2.0;
  }
  cd E = c->ecore + e1 + ej - ek;
  out[0] = E * O * W.orthoFac;
  out[1] = O * W.orthoFac;
  if (e1_out) *e1_out = e1;
  if (ej_out) *ej_out = ej;
  if (ek_out) *ek_out = ek;
}

// psi <- sum_{k=0..5} V^k psi / k!   (applyExp, DQMC/DQMCWalker.cpp:424-431)
void apply_exp(const MC& V, MC& det) {
  MC temp = det;
  for (int i = 1; i < 6; ++i) {
    temp = mul(V, temp);
    for (auto& v : temp.a) v /= (double)i;
    for (size_t q = 0; q < det.a.size(); ++q) det.a[q] += temp.a[q];
  }
}

MC real_times(const MD& E, const MC& X) { return mul(E, X); }

// DQMCWalker::propagatePhaselessRG (DQMC/DQMCWalker.cpp:441-503)
double propagate_walker(const orc_ctx* c, Walker& W, const double* x, double eshift,
                        double* vhs_out, double* fb_out) {
  const int N = c->N, G = c->G;
  std::vector<cd> fb;
  force_bias(c, W, fb);
  if (fb_out)
    for (int g = 0; g < G; ++g) { fb_out[2 * g] = fb[g].real(); fb_out[2 * g + 1] = fb[g].imag(); }
  const double sq = std::sqrt(c->dt);
  cd shift(0.0, 0.0), fbTerm(0.0, 0.0);
  std::vector<cd> fields(G);
  for (int g = 0; g < G; ++g) {
    const cd fieldShift = -sq * (cd(0.0, 1.0) * fb[g] - c->mf[g]);
    const cd fld = x[g] - fieldShift;
    if (c->ref_compat) {
      const std::complex<float> f32(fld);
      fields[g] = cd(f32.real(), f32.imag());
    } else {
      fields[g] = fld;
    }
    shift += (x[g] - fieldShift) * c->mf[g];
    fbTerm += (x[g] * fieldShift - fieldShift * fieldShift / 2.0);
  }
  // VHS: V = i sqrt(dt) sum_g field_g L_g  (DQMC/DQMCWalker.cpp:406-421,459-462); with
  // unrestricted integrals one per spin from that spin's vectors (:525-530)
  auto vhs = [&](int spin) {
    MC V(N, N);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= i; ++j) {
        double pr = 0.0, pi = 0.0;
        for (int g = 0; g < G; ++g) {
          const double l = c->ref_compat ? c->Lfval_s(spin, g, i, j) : c->Lval_s(spin, g, i, j);
          pr += fields[g].real() * l;
          pi += fields[g].imag() * l;
        }
        if (c->ref_compat) {
          pr = (double)(float)pr;
          pi = (double)(float)pi;
        }
        const cd v = cd(-pi, pr) * sq;
        V(i, j) = v;
        V(j, i) = v;
      }
    return V;
  };
  const MC V = vhs(0);
  if (c->uints) {
    // propagatePhaselessU (DQMC/DQMCWalker.cpp:532-540): det[s] <- E_s exp(V_s) E_s det[s]
    const MC Vd = vhs(1);
    W.det[0] = real_times(c->expOB, W.det[0]);
    apply_exp(V, W.det[0]);
    W.det[0] = real_times(c->expOB, W.det[0]);
    W.det[1] = real_times(c->expOBd, W.det[1]);
    apply_exp(Vd, W.det[1]);
    W.det[1] = real_times(c->expOBd, W.det[1]);
  }
  if (vhs_out)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        vhs_out[2 * ((size_t)i * N + j)] = V(i, j).real();
        vhs_out[2 * ((size_t)i * N + j) + 1] = V(i, j).imag();
      }
  if (!c->uints) {
    W.det[0] = real_times(c->expOB, W.det[0]);
    apply_exp(V, W.det[0]);
    W.det[0] = real_times(c->expOB, W.det[0]);
    if (c->walker_rhf) {
      W.det[1] = W.det[0];
    } else {
      W.det[1] = real_times(c->expOB, W.det[1]);
      apply_exp(V, W.det[1]);
      W.det[1] = real_times(c->expOB, W.det[1]);
    }
  }
  const cd oldOverlap = W.trialOverlap;
  const cd newOverlap = overlap(c, W);
  W.trialOverlap = newOverlap;
  const cd imp = std::exp(-sq * shift + fbTerm + c->dt * (eshift + c->propConstant)) *
                 newOverlap / oldOverlap;
  const double theta = std::arg(std::exp(-sq * shift) * newOverlap / oldOverlap);
  double f = std::abs(imp) * std::cos(theta);
  if (f < 1.e-3 || f > 100. || std::isnan(f)) f = 0.;
  return f;
}

// DQMCWalker::orthogonalize (DQMC/DQMCWalker.cpp:295-326), phaseless branch
void orthogonalize_walker(const orc_ctx* c, Walker& W) {
  cd fac = householder_thin_q(W.det[0]);
  if (c->walker_rhf) {
    W.trialOverlap /= (fac * fac);
    W.det[1] = W.det[0];  // det[1] mirrors det[0] for an RHF walker
  } else {
    fac *= householder_thin_q(W.det[1]);
    W.trialOverlap /= fac;
  }
  W.orthoFac = 1.0;
}

void load_dets(const orc_ctx* c, Walker& W, const double* serial) {
  const int N = c->N;
  W.det[0] = MC(N, c->na);
  W.det[1] = MC(N, c->nb);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < c->na; ++j) {
      const size_t q = (size_t)i * c->na + j;
      W.det[0](i, j) = cd(serial[2 * q], serial[2 * q + 1]);
    }
  const size_t off = (size_t)N * c->na;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < c->nb; ++j) {
      const size_t q = off + (size_t)i * c->nb + j;
      W.det[1](i, j) = cd(serial[2 * q], serial[2 * q + 1]);
    }
}

void store_dets(const orc_ctx* c, const Walker& W, double* serial) {
  const int N = c->N;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < c->na; ++j) {
      const size_t q = (size_t)i * c->na + j;
      serial[2 * q] = W.det[0](i, j).real();
      serial[2 * q + 1] = W.det[0](i, j).imag();
    }
  const size_t off = (size_t)N * c->na;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < c->nb; ++j) {
      const size_t q = off + (size_t)i * c->nb + j;
      serial[2 * q] = W.det[1](i, j).real();
      serial[2 * q + 1] = W.det[1](i, j).imag();
    }
}

MD read_rowmajor(const double* p, int r, int cc) {
  MD m(r, cc);
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < cc; ++j) m(i, j) = p[(size_t)i * cc + j];
  return m;
}

}  // namespace

extern "C" {

orc_ctx* orc_create(const afq_ham* ham, const afq_trial* left, int walker_rhf, int nwalk,
                    int ref_compat) {
  orc_ctx* c = new orc_ctx();
  c->N = ham->norb;
  c->na = ham->nalpha;
  c->nb = ham->nbeta;
  c->G = ham->nchol;
  c->GE = ham->nchol;
  c->ecore = ham->ecore;
  const int N = c->N;
  c->h1 = read_rowmajor(ham->h1, N, N);
  c->h1mod = read_rowmajor(ham->h1_mod, N, N);
  c->L.assign(ham->chol, ham->chol + (size_t)c->G * N * N);
  c->uints = ham->h1_dn && ham->h1_mod_dn && ham->chol_dn;
  if (c->uints) {
    c->h1d = read_rowmajor(ham->h1_dn, N, N);
    c->h1modd = read_rowmajor(ham->h1_mod_dn, N, N);
    c->Ld.assign(ham->chol_dn, ham->chol_dn + (size_t)c->G * N * N);
  }
  // floattenCholesky (DQMC/Hamiltonian.cpp:179-220)
  const size_t T = (size_t)N * (N + 1) / 2;
  c->Lf.resize((size_t)c->G * T);
  for (int g = 0; g < c->G; ++g) {
    size_t cnt = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= i; ++j) c->Lf[(size_t)g * T + cnt++] = (float)c->Lval(g, i, j);
  }
  if (c->uints) {  // the down-spin half of floatChol (DQMC/Hamiltonian.cpp:193-196)
    c->Lfd.resize((size_t)c->G * T);
    for (int g = 0; g < c->G; ++g) {
      size_t cnt = 0;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) c->Lfd[(size_t)g * T + cnt++] = (float)c->Lval_s(1, g, i, j);
    }
  }
  c->trial_rhf = left->rhf != 0;
  c->walker_rhf = walker_rhf != 0;
  c->ref_compat = ref_compat != 0;
  c->phi[0] = read_rowmajor(left->phi_up, N, c->na);
  c->phi[1] = c->trial_rhf ? c->phi[0] : read_rowmajor(left->phi_dn, N, c->nb);
  for (int s = 0; s < 2; ++s) {
    c->phiT[s] = transpose(c->phi[s]);
    // rotateCholesky (DQMC/Hamiltonian.cpp:52-133): R_g = Phi^T L_g, col-major n x N
    // (unrestricted integrals: spin s's vectors, :99-107)
    const int n = c->phi[s].c;
    c->R[s].assign((size_t)c->G * n * N, 0.0);
    for (int g = 0; g < c->G; ++g)
      for (int m = 0; m < N; ++m) {
        const double* prow = &c->phiT[s].a[0];  // phiT(i, m) at i + m*n
        for (int k = 0; k < N; ++k) {
          const double l = c->Lval_s(s, g, m, k);
          if (l == 0.0) continue;
          double* r = &c->R[s][(size_t)g * n * N + (size_t)k * n];
          for (int i = 0; i < n; ++i) r[i] += prow[i + (size_t)m * n] * l;
        }
      }
  }
  build_slots(c);
  c->walkers.resize(nwalk);
  c->weights.assign(nwalk, 0.0);
  for (auto& W : c->walkers) {
    W.det[0] = MC(N, c->na);
    W.det[1] = MC(N, c->nb);
  }
  return c;
}

void orc_destroy(orc_ctx* c) { delete c; }

void orc_set_nchol_ene(orc_ctx* c, int n) { c->GE = n; }

// DQMCWalker::prepPropR (DQMC/DQMCWalker.cpp:32-103), phaseless
void orc_prep_prop(orc_ctx* c, const double* ref_up, const double* ref_dn,
                   const double* rdm_up, const double* rdm_dn, double dt, double ene0) {
  const int N = c->N;
  c->dt = dt;
  MD green[2];
  if (rdm_up && rdm_dn) {
    green[0] = read_rowmajor(rdm_up, N, N);
    green[1] = read_rowmajor(rdm_dn, N, N);
  } else {
    const double* refs[2] = {ref_up, ref_dn};
    const int ns[2] = {c->na, c->nb};
    for (int s = 0; s < 2; ++s) {
      MD P = read_rowmajor(refs[s], N, ns[s]);
      // G = P (P^T P)^-1 P^T
      MC PtP(ns[s], ns[s]);
      for (int i = 0; i < ns[s]; ++i)
        for (int j = 0; j < ns[s]; ++j) {
          double a = 0;
          for (int k = 0; k < N; ++k) a += P(k, i) * P(k, j);
          PtP(i, j) = a;
        }
      std::vector<int> piv;
      int nsw;
      lu_factor(PtP, piv, nsw);
      MC inv = lu_inverse(PtP, piv);
      green[s] = MD(N, N);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          double a = 0;
          for (int p = 0; p < ns[s]; ++p)
            for (int q = 0; q < ns[s]; ++q) a += P(i, p) * inv(p, q).real() * P(j, q);
          green[s](i, j) = a;
        }
    }
  }
  // restricted: one h1' from h1Mod; unrestricted (prepPropU, DQMC/DQMCWalker.cpp:141-180):
  // one per spin from that spin's h1Mod and vectors, the shifts summing both spins
  MD h1p[2] = {c->h1mod, c->uints ? c->h1modd : c->h1mod};
  auto lv = [&](int spin, int g, int i, int j) {
    return c->ref_compat ? c->Lfval_s(spin, g, i, j) : c->Lval_s(spin, g, i, j);
  };
  cd constant = ene0 - c->ecore;
  c->mf.assign(c->G, cd(0.0));
  for (int g = 0; g < c->G; ++g) {
    double a = 0.0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        a += c->uints ? green[0](i, j) * lv(0, g, i, j) + green[1](i, j) * lv(1, g, i, j)
                      : (green[0](i, j) + green[1](i, j)) * lv(0, g, i, j);
    const cd mfs = cd(0.0, a);  // sum G o (i L)
    constant -= mfs * mfs / 2.0;
    for (int spin = 0; spin < (c->uints ? 2 : 1); ++spin)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) h1p[spin](i, j) -= (mfs * cd(0.0, lv(spin, g, i, j))).real();  // mf * i L is real
    c->mf[g] = mfs;
  }
  c->propConstant = constant - ene0;
  c->expOB = expm_sym(h1p[0], -dt / 2.0);
  if (c->uints) c->expOBd = expm_sym(h1p[1], -dt / 2.0);
  c->prepped = true;
}

void orc_get_prop_data(orc_ctx* c, double* mf, double* expob, double* pc) {
  if (mf)
    for (int g = 0; g < c->G; ++g) { mf[2 * g] = c->mf[g].real(); mf[2 * g + 1] = c->mf[g].imag(); }
  if (expob)
    for (int i = 0; i < c->N; ++i)
      for (int j = 0; j < c->N; ++j) expob[(size_t)i * c->N + j] = c->expOB(i, j);
  if (pc) { pc[0] = c->propConstant.real(); pc[1] = c->propConstant.imag(); }
}

void orc_set_walkers(orc_ctx* c, const double* dets, const double* overlaps,
                     const double* weights) {
  const size_t ms = (size_t)c->N * (c->na + c->nb);
  for (size_t w = 0; w < c->walkers.size(); ++w) {
    Walker& W = c->walkers[w];
    load_dets(c, W, dets + 2 * ms * w);
    W.orthoFac = 1.0;
    if (overlaps) W.trialOverlap = cd(overlaps[2 * w], overlaps[2 * w + 1]);
    else W.trialOverlap = overlap(c, W);
    c->weights[w] = weights ? weights[w] : 1.0;
  }
}

void orc_get_walkers(orc_ctx* c, double* dets, double* overlaps, double* weights) {
  const size_t ms = (size_t)c->N * (c->na + c->nb);
  for (size_t w = 0; w < c->walkers.size(); ++w) {
    const Walker& W = c->walkers[w];
    if (dets) store_dets(c, W, dets + 2 * ms * w);
    if (overlaps) { overlaps[2 * w] = W.trialOverlap.real(); overlaps[2 * w + 1] = W.trialOverlap.imag(); }
    if (weights) weights[w] = c->weights[w];
  }
}

double orc_propagate_walker(orc_ctx* c, int w, const double* x, double eshift,
                            double* vhs_out, double* fb_out) {
  return propagate_walker(c, c->walkers[w], x, eshift, vhs_out, fb_out);
}

void orc_propagate(orc_ctx* c, const double* draws, double eshift, double weight_cap,
                   double* factors, int64_t* n_capped) {
  int64_t ncap = 0;
  for (size_t w = 0; w < c->walkers.size(); ++w) {
    double f = 1.0;
    if (c->weights[w] > 1.e-8) {
      f = propagate_walker(c, c->walkers[w], draws + (size_t)w * c->G, eshift, nullptr, nullptr);
      c->weights[w] *= f;
    }
    if (factors) factors[w] = f;
    if (c->weights[w] > weight_cap) { c->weights[w] = 0.0; ++ncap; }
  }
  if (n_capped) *n_capped = ncap;
}

void orc_orthogonalize(orc_ctx* c) {
  for (auto& W : c->walkers) orthogonalize_walker(c, W);
}
void orc_orthogonalize_walker(orc_ctx* c, int w) { orthogonalize_walker(c, c->walkers[w]); }

void orc_ham_and_overlap_walker(orc_ctx* c, int w, double* out4) {
  cd o[2];
  ham_and_overlap(c, c->walkers[w], c->GE, o);
  out4[0] = o[0].real(); out4[1] = o[0].imag(); out4[2] = o[1].real(); out4[3] = o[1].imag();
}

void orc_energy_parts_walker(orc_ctx* c, int w, double* e1, double* ej, double* ek) {
  cd o[2], a, b, d;
  ham_and_overlap(c, c->walkers[w], c->GE, o, &a, &b, &d);
  e1[0] = a.real(); e1[1] = a.imag();
  ej[0] = b.real(); ej[1] = b.imag();
  ek[0] = d.real(); ek[1] = d.imag();
}

void orc_force_bias_walker(orc_ctx* c, int w, double* fb) {
  std::vector<cd> v;
  force_bias(c, c->walkers[w], v);
  for (int g = 0; g < c->G; ++g) { fb[2 * g] = v[g].real(); fb[2 * g + 1] = v[g].imag(); }
}

void orc_overlap_walker(orc_ctx* c, int w, double* ovlp) {
  cd o = overlap(c, c->walkers[w]);
  ovlp[0] = o.real(); ovlp[1] = o.imag();
}

void orc_green_walker(orc_ctx* c, int w, int slot, double* theta, double* green) {
  const Slot& s = c->slots[slot];
  MC th;
  green_slot(c, c->walkers[w], s, &th, nullptr);
  const int N = c->N;
  if (theta)
    for (int k = 0; k < N; ++k)
      for (int i = 0; i < s.n; ++i) {
        theta[2 * ((size_t)k * s.n + i)] = th(k, i).real();
        theta[2 * ((size_t)k * s.n + i) + 1] = th(k, i).imag();
      }
  if (green)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        cd g(0.0, 0.0);
        for (int k = 0; k < s.n; ++k) g += th(i, k) * c->phi[s.phi](j, k);
        green[2 * ((size_t)i * N + j)] = g.real();
        green[2 * ((size_t)i * N + j) + 1] = g.imag();
      }
}

// One walker's term of the mixed-estimator 1-RDM (DQMC/MixedEstimator.cpp:492-510):
// rdm = (rdm cw + w Re(sample)) / (cw + w), cw += w, where sample is
// RHF::oneRDM's (psi_s (Phi^T psi_s)^-1 Phi^T)^T summed over the spins
// (DQMC/RHF.cpp:64-67; spin_summed, one N x N block) or UHF::oneRDM's per-spin
// pair (DQMC/UHF.cpp:68-72; two blocks, up then down).  rdm is row-major.
static void one_rdm_walker(const orc_ctx* c, const Walker& W, double w, int spin_summed, double* rdm,
                           double* cw) {
  const int N = c->N;
  const size_t NN = (size_t)N * N;
  const int nblk = spin_summed ? 1 : 2;
  std::vector<double> sample(nblk * NN, 0.0);
  for (size_t q = 0; q < c->slots.size(); ++q) {  // slot 0: up, slot 1: down
    const Slot& s = c->slots[q];
    MC th;
    green_slot(c, W, s, &th, nullptr);
    MC G(N, N);  // psi (Phi^T psi)^-1 Phi^T
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < s.n; ++k) {
        const double pj = c->phi[s.phi](j, k);
        for (int i = 0; i < N; ++i) G(i, j) += th(i, k) * pj;
      }
    for (int b = 0; b < nblk; ++b) {
      // an RHF walker against an RHF trial: one slot of multiplicity 2 for both spins
      const double m = spin_summed ? s.mult : (((int)q == b || s.mult == 2) ? 1.0 : 0.0);
      if (m == 0.0) continue;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) sample[b * NN + (size_t)i * N + j] += m * G(j, i).real();
    }
  }
  for (size_t q = 0; q < nblk * NN; ++q) rdm[q] *= *cw;
  for (size_t q = 0; q < nblk * NN; ++q) rdm[q] += w * sample[q];
  *cw += w;
  for (size_t q = 0; q < nblk * NN; ++q) rdm[q] /= *cw;
}

void orc_one_rdm(orc_ctx* c, int spin_summed, double* rdm, double* cw) {
  for (size_t w = 0; w < c->walkers.size(); ++w)
    if (c->weights[w] != 0.) one_rdm_walker(c, c->walkers[w], c->weights[w], spin_summed, rdm, cw);
}

void orc_ham_and_overlap_dets(orc_ctx* c, const double* dets, int ndet, int nchol_ene,
                              double* out) {
  const size_t ms = (size_t)c->N * (c->na + c->nb);
  for (int d = 0; d < ndet; ++d) {
    Walker W;
    load_dets(c, W, dets + 2 * ms * d);
    cd o[2];
    ham_and_overlap(c, W, nchol_ene, o);
    out[4 * d] = o[0].real(); out[4 * d + 1] = o[0].imag();
    out[4 * d + 2] = o[1].real(); out[4 * d + 3] = o[1].imag();
  }
}

// comb reconfiguration (DQMC/MixedEstimator.cpp:595-622), one rank
void orc_comb(orc_ctx* c, double zeta, double* total_weight) {
  const size_t W = c->walkers.size();
  std::vector<double> cum(W);
  cum[0] = c->weights[0];
  for (size_t i = 1; i < W; ++i) cum[i] = cum[i - 1] + c->weights[i];
  const double tot = cum[W - 1];
  std::vector<MC> d0(W), d1(W);
  std::vector<cd> ov(W);
  for (size_t w = 0; w < W; ++w) {
    const double z = (w + zeta) / W;
    const size_t idx = std::lower_bound(cum.begin(), cum.end(), z * tot) - cum.begin();
    d0[w] = c->walkers[idx].det[0];
    d1[w] = c->walkers[idx].det[1];
    ov[w] = c->walkers[idx].trialOverlap;
  }
  for (size_t w = 0; w < W; ++w) {
    c->walkers[w].det[0] = d0[w];
    c->walkers[w].det[1] = d1[w];
    c->walkers[w].trialOverlap = ov[w];
    c->weights[w] = tot / W;
  }
  if (total_weight) *total_weight = tot;
}

void orc_blocking(const double* weights, const double* energies, int n, double* out2) {
  static const int sizes[] = {1, 2, 5, 10, 20, 50, 70, 100, 200, 500};
  double wsum = 0, wesum = 0;
  for (int i = 0; i < n; ++i) { wsum += weights[i]; wesum += weights[i] * energies[i]; }
  out2[0] = wesum / wsum;
  double prevError = 0., plateau = -1.;
  for (int b : sizes) {
    if (b > n / 2.) break;
    const int nb = n / b;
    std::vector<double> bw(nb, 0.0), be(nb, 0.0);
    for (int j = 0; j < nb; ++j) {
      double sw = 0, se = 0;
      for (int q = j * b; q < (j + 1) * b; ++q) { sw += weights[q]; se += weights[q] * energies[q]; }
      bw[j] = sw;
      be[j] = se / sw;
    }
    double v1 = 0, v2 = 0, mean = 0;
    for (int j = 0; j < nb; ++j) { v1 += bw[j]; v2 += bw[j] * bw[j]; mean += bw[j] * be[j]; }
    mean /= v1;
    double acc = 0;
    for (int j = 0; j < nb; ++j) acc += bw[j] * (be[j] - mean) * (be[j] - mean);
    const double err = std::sqrt(acc / (v1 - v2 / v1) / (nb - 1));
    if (err < 1.05 * prevError && plateau < 0.) plateau = std::max(err, prevError);
    prevError = err;
  }
  out2[1] = plateau;
}

// calcMixedEstimatorLongProp (DQMC/MixedEstimator.cpp:309-699) with nranks
// virtual ranks; rank r owns walkers [r*nwalk, (r+1)*nwalk) and generator
// std::mt19937(seed + r) (executables/DQMC.cpp:58).
int orc_run_long_prop(const afq_ham* ham, const afq_trial* left, const double* ref_up,
                      const double* ref_dn, const double* init_up, const double* init_dn,
                      const orc_run_params* p, double* weights_out, double* energies_out,
                      orc_run_result* res) {
  const int R = p->nranks, nw = p->nwalk, W = R * nw;
  orc_ctx* c = orc_create(ham, left, p->walker_rhf, W, p->ref_compat);
  const int N = c->N;
  const size_t ms = (size_t)N * (c->na + c->nb);
  const double dt = p->dt;
  // reference determinant in serial layout
  std::vector<double> ref(2 * ms, 0.0), init(2 * ms, 0.0);
  for (int i = 0; i < N; ++i) {
    for (int j = 0; j < c->na; ++j) {
      ref[2 * ((size_t)i * c->na + j)] = ref_up[(size_t)i * c->na + j];
      init[2 * ((size_t)i * c->na + j)] = init_up[(size_t)i * c->na + j];
    }
    for (int j = 0; j < c->nb; ++j) {
      const size_t q = (size_t)N * c->na + (size_t)i * c->nb + j;
      ref[2 * q] = ref_dn[(size_t)i * c->nb + j];
      init[2 * q] = init_dn[(size_t)i * c->nb + j];
    }
  }
  // reference energy and Cholesky truncation (:352-384)
  double o4[4];
  orc_ham_and_overlap_dets(c, ref.data(), 1, c->G, o4);
  const cd refEnergy = cd(o4[0], o4[1]) / cd(o4[2], o4[3]);
  const double ene0 = (p->ene0_guess == 1.e10) ? refEnergy.real() : p->ene0_guess;
  int nchol = c->G;
  cd delta(0.0, 0.0);
  const int G = c->G;
  const int cand[5] = {int(0.3 * G), int(0.4 * G), int(0.5 * G), int(0.6 * G), int(0.7 * G)};
  for (int i = 0; i < 5; ++i) {
    orc_ham_and_overlap_dets(c, ref.data(), 1, cand[i], o4);
    const cd e = cd(o4[0], o4[1]) / cd(o4[2], o4[3]);
    if (std::abs(refEnergy - e) < p->cholesky_threshold) {
      nchol = cand[i];
      delta = refEnergy - e;
      break;
    }
  }
  c->GE = nchol;
  // prepProp with the reference dets (:395-396)
  orc_prep_prop(c, ref_up, ref_dn, nullptr, nullptr, dt, ene0);
  // walkers start from the right trial (:398-416)
  std::vector<double> dets((size_t)W * 2 * ms);
  for (int w = 0; w < W; ++w) std::memcpy(&dets[(size_t)w * 2 * ms], init.data(), 2 * ms * sizeof(double));
  orc_set_walkers(c, dets.data(), nullptr, nullptr);
  std::vector<double> eloc(W);
  const int nsweeps = p->nsweeps, nsteps = p->nsteps;
  std::vector<double> totalWeights(nsweeps, 0.0), totalEnergies(nsweeps, 0.0);
  double esum = 0;
  for (int w = 0; w < W; ++w) {
    cd o[2];
    ham_and_overlap(c, c->walkers[w], c->GE, o);
    eloc[w] = (o[0] / o[1]).real();
    esum += eloc[w];
  }
  totalWeights[0] = W;
  totalEnergies[0] = esum / totalWeights[0] + delta.real();
  std::vector<std::mt19937> gens;
  for (int r = 0; r < R; ++r) gens.emplace_back((unsigned)(p->seed + r));
  double eshift = totalEnergies[0];
  double totalWeight = W;
  std::vector<double> averageEnergy(R, totalEnergies[0]), averageEnergyEql(R, totalEnergies[0]);
  double averageNum = 0., averageDenom = 0., averageNumEql = 0., averageDenomEql = 0.;
  double eEstimate = totalEnergies[0];
  int64_t nLarge = 0;
  const double weightCap = p->weight_cap > 0 ? p->weight_cap : std::max(100., nw / 10.);
  std::vector<double> x(G);
  double propTime = 0, eneTime = 0;
  // mixed-estimator 1-RDM (:440-450, 492-510): RHF trial spin-summed, UHF trial per spin
  const int rdm_summed = c->trial_rhf ? 1 : 0;
  double rdm_cw = 0.0;
  if (p->one_rdm_out) std::fill(p->one_rdm_out, p->one_rdm_out + (rdm_summed ? 1 : 2) * (size_t)N * N, 0.0);
  const int64_t last = (int64_t)nsweeps * nsteps;
  for (int64_t step = 1; step < last; ++step) {
    if (p->max_steps > 0 && step > p->max_steps) break;
    for (int r = 0; r < R; ++r)
      if (step * dt < 10.) averageEnergy[r] = averageEnergyEql[r];
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < R; ++r)
      for (int wl = 0; wl < nw; ++wl) {
        const int w = r * nw + wl;
        Walker& Wk = c->walkers[w];
        if (c->weights[w] > 1.e-8) {
          for (int g = 0; g < G; ++g) x[g] = Wk.normal(gens[r]);
          c->weights[w] *= propagate_walker(c, Wk, x.data(), eshift, nullptr, nullptr);
        }
        if (c->weights[w] > weightCap) { c->weights[w] = 0.; ++nLarge; }
      }
    auto t1 = std::chrono::steady_clock::now();
    propTime += std::chrono::duration<double>(t1 - t0).count();
    totalWeight = 0;
    for (int w = 0; w < W; ++w) totalWeight += c->weights[w];
    eshift = eEstimate - 0.1 * std::log(totalWeight / (double)W) / dt;
    if (step % p->ortho_steps == 0)
      for (int w = 0; w < W; ++w) orthogonalize_walker(c, c->walkers[w]);
    if (step % nsteps == 0) {
      const int block = (int)(step / nsteps);
      auto t2 = std::chrono::steady_clock::now();
      for (int r = 0; r < R; ++r)
        for (int wl = 0; wl < nw; ++wl) {
          const int w = r * nw + wl;
          if (c->weights[w] != 0.) {
            if (p->one_rdm_out && step * dt > 10. && step > (int64_t)p->burn_iter * nsteps)
              one_rdm_walker(c, c->walkers[w], c->weights[w], rdm_summed, p->one_rdm_out, &rdm_cw);
            cd o[2];
            ham_and_overlap(c, c->walkers[w], c->GE, o);
            eloc[w] = (o[0] / o[1]).real() + delta.real();
            if (std::isnan(eloc[w]) || std::isinf(eloc[w])) {
              delete c;
              return AFQ_ERR_NUMERIC;
            } else if (std::abs(eloc[w] - averageEnergy[r]) > std::sqrt(2. / dt)) {
              ++nLarge;
              eloc[w] = eloc[w] > averageEnergy[r] ? averageEnergy[r] + std::sqrt(2. / dt)
                                                   : averageEnergy[r] - std::sqrt(2. / dt);
            }
          } else {
            eloc[w] = 0.;
          }
        }
      eneTime += std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
      double we = 0;
      for (int w = 0; w < W; ++w) we += eloc[w] * c->weights[w];
      totalWeights[block] = totalWeight;
      totalEnergies[block] = we / totalWeight;
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
      // comb on rank 0 with rank 0's generator (:595-622)
      std::uniform_real_distribution<double> uniform(0., 1.);
      const double zeta = uniform(gens[0]);
      double tot;
      orc_comb(c, zeta, &tot);
      totalWeight = tot;
    }
  }
  for (int i = 0; i < nsweeps; ++i) {
    if (weights_out) weights_out[i] = totalWeights[i];
    if (energies_out) energies_out[i] = totalEnergies[i];
  }
  if (res) {
    res->ref_energy = refEnergy.real();
    res->nchol_ene = nchol;
    res->delta = delta.real();
    res->n_large_deviations = nLarge;
    res->prop_time = propTime;
    res->ene_time = eneTime;
    res->one_rdm_weight = rdm_cw;
  }
  delete c;
  return AFQ_OK;
}

double orc_bench(orc_ctx* c, int nthreads, int nsteps, int energy_every, int ortho_steps,
                 uint64_t seed, double* seconds) {
  const int W = (int)c->walkers.size();
  if (nthreads < 1) nthreads = 1;
  std::atomic<long> done{0};
  auto work = [&](int t) {
    std::mt19937 gen((unsigned)(seed + t));
    std::normal_distribution<double> normal(0., 1.);
    std::vector<double> x(c->G);
    for (int step = 1; step <= nsteps; ++step)
      for (int w = t; w < W; w += nthreads) {
        Walker& Wk = c->walkers[w];
        for (int g = 0; g < c->G; ++g) x[g] = normal(gen);
        const double f = propagate_walker(c, Wk, x.data(), 0.0, nullptr, nullptr);
        (void)f;
        if (ortho_steps > 0 && step % ortho_steps == 0) orthogonalize_walker(c, Wk);
        if (energy_every > 0 && step % energy_every == 0) {
          cd o[2];
          ham_and_overlap(c, Wk, c->GE, o);
        }
        done.fetch_add(1);
      }
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
  for (auto& x : th) x.join();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (seconds) *seconds = s;
  return (double)done.load() / s;
}

void orc_normals(uint64_t seed, int n, double* out) {
  std::mt19937 gen((unsigned)seed);
  std::normal_distribution<double> normal(0., 1.);
  for (int i = 0; i < n; ++i) out[i] = normal(gen);
}

uint32_t orc_mt19937_10000th(void) {
  std::mt19937 gen;
  uint32_t v = 0;
  for (int i = 0; i < 10000; ++i) v = gen();
  return v;
}

}  // extern "C"
