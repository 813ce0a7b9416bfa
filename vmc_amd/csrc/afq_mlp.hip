// Learned exchange-energy surrogate: a GELU MLP F -> H -> H -> H -> 1 with
// FP32 master weights, every GEMM on BF16 MFMA (operands rounded to bf16 when
// staged in LDS) with FP32 accumulation, trained with MSE + Adam on the GPU.
//
// Input (per walker): Re fb_g (g < G) and Re diag G_s (two spins, N each),
// standardised with the training set's mean / std.  Target: the exchange energy
// E_K = 1/2 sum_s m_s sum_g tr(f_{s,g} f_{s,g}) relative to the Coulomb energy
// E_J = 1/2 sum_g fb_g^2 that is computed exactly anyway: y = E_K - E_J.  The two
// are strongly correlated and largely cancel in E_L, so y has a much smaller
// spread than E_K.
#include <math.h>

#include <stdexcept>
#include <string>

#include "afq_device.hpp"
#include "afq_internal.hpp"

namespace afq {

static inline unsigned cdivu(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------- GEMM (bf16 operands, fp32 accumulate)
template <class LA, class LB, class EP>
__global__ __launch_bounds__(256) void k_gemm_bf16e(LA la, LB lb, EP ep, int M, int N, int K) {
  using Eng = MmaBF16E<64, 64, 2, 2>;
  __shared__ __attribute__((aligned(16))) u16 lds[Eng::LDS_BYTES / 2];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  Eng e;
  e.run([&](int r, int k) { return (r0 + r < M && k < K) ? la(r0 + r, k) : 0.f; },
        [&](int k, int c) { return (k < K && c0 + c < N) ? lb(k, c0 + c) : 0.f; }, K, lds);
  e.store([&](int r, int c, float v) {
    if (r0 + r < M && c0 + c < N) ep(r0 + r, c0 + c, v);
  });
}
template <class LA, class LB, class EP>
static void gemm(const LA& la, const LB& lb, const EP& ep, int M, int N, int K, hipStream_t st) {
  hipLaunchKernelGGL((k_gemm_bf16e<LA, LB, EP>), dim3(cdivu(N, 64), cdivu(M, 64)), dim3(256), 0, st, la, lb, ep, M,
                     N, K);
}

// operand views
struct InRows {  // standardised network input rows, gathered by idx
  const float* X; const int* idx; const float* mu; const float* sd; int F;
  __device__ float operator()(int r, int k) const {
    const int row = idx ? idx[r] : r;
    return (X[(int64_t)row * F + k] - mu[k]) / sd[k];
  }
};
struct InRowsT {  // transpose of InRows: (feature, batch)
  InRows a;
  __device__ float operator()(int r, int k) const { return a(k, r); }
};
struct Rows {  // H[r][k], row length K
  const float* H; int K;
  __device__ float operator()(int r, int k) const { return H[(int64_t)r * K + k]; }
};
struct RowsT {  // (r, k) -> H[k][r]
  const float* H; int K;
  __device__ float operator()(int r, int k) const { return H[(int64_t)k * K + r]; }
};
struct DZ {  // dZ[b][c] = dH[b][c] * gelu'(z[b][c]) (or dH itself for the output layer)
  const float* dH; const float* z; int N; int act;
  __device__ float operator()(int b, int c) const {
    const float d = dH[(int64_t)b * N + c];
    return act ? d * gelu_d(z[(int64_t)b * N + c]) : d;
  }
};
struct DZT {  // (c, b) view of DZ
  DZ a;
  __device__ float operator()(int r, int k) const { return a(k, r); }
};
struct WB {  // W[k][c] (in x out)
  const float* W; int N;
  __device__ float operator()(int k, int c) const { return W[(int64_t)k * N + c]; }
};
struct WBT {  // (o, i) -> W[i][o]
  const float* W; int N;
  __device__ float operator()(int k, int c) const { return W[(int64_t)c * N + k]; }
};
struct FwdEp {
  float* z; float* h; const float* bias; int N; int act;
  __device__ void operator()(int r, int c, float v) const {
    const float zz = v + bias[c];
    z[(int64_t)r * N + c] = zz;
    h[(int64_t)r * N + c] = act ? gelu(zz) : zz;
  }
};
struct StoreEp {
  float* out; int N;
  __device__ void operator()(int r, int c, float v) const { out[(int64_t)r * N + c] = v; }
};

// ---------------------------------------------------------------- small kernels
__global__ void k_colsum_dz(DZ dz, float* db, int B, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += dz(b, c);
  db[c] = s;
}
// dy = (y - t_std)/B with t_std = (t - t_mu)/t_sd ; loss = 1/2 mean (y - t_std)^2
__global__ void k_loss(const float* y, const float* t, const int* idx, float* dy, float* lossp, int B, float tmu,
                       float tsd) {
  __shared__ float red[256];
  float l = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float ts = (t[idx ? idx[b] : b] - tmu) / tsd;
    const float d = y[b] - ts;
    dy[b] = d / B;
    l += 0.5f * d * d;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[0] = red[0] / B;
}
__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// features of walkers idx[0..n): row j of X at dst0 + j
__global__ __launch_bounds__(256) void k_features(const double* fb, const double* theta, const double* phi0,
                                                  const double* phi1, const int* idx, float* X, int64_t dst0, int G,
                                                  int N, int64_t Ktot, Slot s0, Slot s1, int nslots, int F) {
  const int j = blockIdx.x;
  const int w = idx ? idx[j] : j;
  float* x = X + (dst0 + j) * F;
  for (int g = threadIdx.x; g < G; g += blockDim.x) x[g] = (float)fb[(int64_t)w * 2 * G + 2 * g];
  for (int e = threadIdx.x; e < 2 * N; e += blockDim.x) {
    const int sp = e / N, i = e - sp * N;
    const Slot sl = (sp == 0 || nslots == 1) ? s0 : s1;
    const double* ph = sl.phi == 0 ? phi0 : phi1;
    const double* th = theta + (int64_t)w * 2 * Ktot + 2 * sl.koff + (int64_t)i * 2 * sl.n;
    double d = 0.0;
    for (int q = 0; q < sl.n; ++q) d += th[2 * q] * ph[(int64_t)i * sl.n + q];
    x[G + e] = (float)d;
  }
}

// exact exchange energy E_K of walkers idx[0..n) from the exchange partials; yout
// (training target) receives E_K - E_J, E_J = 1/2 sum_{g<GE} Re fb_g^2
__global__ void k_ek(const double* Xp, const double* fb, const int* idx, int n, double* ek, float* yout,
                     int64_t dst0, int W, int nslots, int m0, int m1, int ngt0, int ngt1, int ngt_cap, int G, int GE) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int w = idx ? idx[j] : j;
  double s = 0.0;
  for (int q = 0; q < nslots; ++q) {
    const int m = q == 0 ? m0 : m1, nt = q == 0 ? ngt0 : ngt1;
    for (int t = 0; t < nt; ++t) s += m * Xp[2 * (((int64_t)q * W + w) * ngt_cap + t)];
  }
  s *= 0.5;
  if (ek) ek[w] = s;
  if (yout) {
    double ej = 0.0;
    for (int g = 0; g < GE; ++g) {
      const double a = fb[(int64_t)w * 2 * G + 2 * g], b = fb[(int64_t)w * 2 * G + 2 * g + 1];
      ej += a * a - b * b;
    }
    yout[dst0 + j] = (float)(s - 0.5 * ej);
  }
}

// E_J = 1/2 sum_{g<GE} Re fb_g^2 of every walker (risk score of the control variate)
__global__ __launch_bounds__(64) void k_ej(const double* fb, double* ej, int G, int GE) {
  const int w = blockIdx.x, lane = threadIdx.x;
  double jr = 0.0;
  for (int g = lane; g < GE; g += 64) {
    const double a = fb[(int64_t)w * 2 * G + 2 * g], b = fb[(int64_t)w * 2 * G + 2 * g + 1];
    jr += a * a - b * b;
  }
  jr = wave_sum(jr);
  if (lane == 0) ej[w] = 0.5 * jr;
}

// control-variate local energy.  S_w - soff predicts y = E_K - E_J, so the
// surrogate energy is e_s = ecore + e1 - (S_w - soff) + delta and the exact one
// e_x = ecore + e1 - y + delta; both are clipped to center[w] -+ halfw (halfw > 0)
// as the population-control estimator clips every local energy.  A walker in the
// exact set (inclusion probability 1/pinv[w]) reports e_s + pinv[w] (e_x - e_s),
// an unbiased (Horvitz-Thompson) estimate of its clipped exact energy for any soff
// fixed before the exact set is drawn; dres[w] = e_x - e_s for the exact set.
__global__ __launch_bounds__(64) void k_energy_cv(const double* fb, const double* e1, const double* S,
                                                  const double* ek, const float* pinv,
                                                  const double* weights, double* eloc, double* dres, int G, int GE,
                                                  int nslots, int m0, int m1, double ecore, double delta, double soff,
                                                  const double* center, double halfw) {
  const int w = blockIdx.x, lane = threadIdx.x;
  double jr = 0.0;
  for (int g = lane; g < GE; g += 64) {
    const double a = fb[(int64_t)w * 2 * G + 2 * g], b = fb[(int64_t)w * 2 * G + 2 * g + 1];
    jr += a * a - b * b;
  }
  jr = wave_sum(jr);
  if (lane == 0) {
    const double lo = center && halfw > 0.0 ? center[w] - halfw : -INFINITY;
    const double hi = center && halfw > 0.0 ? center[w] + halfw : INFINITY;
    double base = ecore + delta;
    for (int s = 0; s < nslots; ++s) base += (s == 0 ? m0 : m1) * e1[2 * ((int64_t)w * nslots + s)];
    const double es = fmin(fmax(base - (S[w] - soff), lo), hi);
    double er = es, d = 0.0;
    if (pinv[w] > 0.f) {
      const double ex = fmin(fmax(base - (ek[w] - 0.5 * jr), lo), hi);
      d = ex - es;
      er += (double)pinv[w] * d;
    }
    const bool live = weights == nullptr || weights[w] != 0.0;
    eloc[w] = live ? er : 0.0;
    dres[w] = live ? d : 0.0;
  }
}

__global__ void k_pred_to_double(const float* y, double* S, int n, float tmu, float tsd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) S[j] = (double)y[j] * tsd + tmu;
}

// ---------------------------------------------------------------- host side
static void hcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void mlp_forward(const Mlp& m, const float* X, const int* idx, int B, hipStream_t st) {
  const int* d = m.dims;
  gemm(InRows{X, idx, m.mu, m.sd, d[0]}, WB{m.w[0], d[1]}, FwdEp{m.z[0], m.h[0], m.b[0], d[1], 1}, B, d[1], d[0], st);
  for (int l = 1; l < 4; ++l)
    gemm(Rows{m.h[l - 1], d[l]}, WB{m.w[l], d[l + 1]}, FwdEp{m.z[l], m.h[l], m.b[l], d[l + 1], l < 3 ? 1 : 0}, B,
         d[l + 1], d[l], st);
}

// gradients of 1/2 mean (y - t)^2 for the batch (m.h[3] holds y); dy already in m.dh[3]
void mlp_backward(const Mlp& m, const float* X, const int* idx, int B, hipStream_t st) {
  const int* d = m.dims;
  for (int l = 3; l >= 0; --l) {
    const DZ dz{m.dh[l], m.z[l], d[l + 1], l < 3 ? 1 : 0};
    // dW_l = in_l^T dZ_l
    if (l == 0)
      gemm(InRowsT{InRows{X, idx, m.mu, m.sd, d[0]}}, dz, StoreEp{m.gw[0], d[1]}, d[0], d[1], B, st);
    else
      gemm(RowsT{m.h[l - 1], d[l]}, dz, StoreEp{m.gw[l], d[l + 1]}, d[l], d[l + 1], B, st);
    hipLaunchKernelGGL(k_colsum_dz, dim3(cdivu(d[l + 1], 256)), dim3(256), 0, st, dz, m.gb[l], B, d[l + 1]);
    // dH_{l-1} = dZ_l W_l^T
    if (l > 0) gemm(dz, WBT{m.w[l], d[l + 1]}, StoreEp{m.dh[l - 1], d[l]}, B, d[l], d[l + 1], st);
  }
}

float mlp_loss(const Mlp& m, const float* t, const int* idx, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, st, m.h[3], t, idx, m.dh[3], m.lossbuf, B, m.t_mu, m.t_sd);
  float l = 0.f;
  hcheck(hipMemcpyAsync(&l, m.lossbuf, sizeof(float), hipMemcpyDeviceToHost, st), "loss D2H");
  hcheck(hipStreamSynchronize(st), "loss sync");
  return l;
}

void mlp_adam(Mlp& m, float lr, hipStream_t st) {
  m.t += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)m.t), c2 = 1.f - powf(b2, (float)m.t);
  for (int l = 0; l < 4; ++l) {
    const int64_t nw = (int64_t)m.dims[l] * m.dims[l + 1], nb = m.dims[l + 1];
    hipLaunchKernelGGL(k_adam, dim3(cdivu(nw, 256) > 1024 ? 1024 : cdivu(nw, 256)), dim3(256), 0, st, m.w[l], m.gw[l],
                       m.mw[l], m.vw[l], nw, lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(cdivu(nb, 256)), dim3(256), 0, st, m.b[l], m.gb[l], m.mb[l], m.vb[l], nb, lr, b1,
                       b2, eps, c1, c2);
  }
}

void launch_features(const DevState& s, const int* idx, int n, float* X, int64_t dst0, hipStream_t st) {
  hipLaunchKernelGGL(k_features, dim3(n), dim3(256), 0, st, s.fb, s.theta, s.phi[0], s.phi[1], idx, X, dst0, s.G, s.N,
                     s.Ktot, s.slot[0], s.slot[s.nslots > 1 ? 1 : 0], s.nslots, s.G + 2 * s.N);
}

void launch_ek(const DevState& s, const int* idx, int n, double* ek, float* yout, int64_t dst0, hipStream_t st) {
  const int ng0 = (int)cdivu(s.GE, exchange_tile_g(s.slot[0].n));
  const int ng1 = s.nslots > 1 ? (int)cdivu(s.GE, exchange_tile_g(s.slot[1].n)) : 0;
  hipLaunchKernelGGL(k_ek, dim3(cdivu(n, 256)), dim3(256), 0, st, s.Xp, s.fb, idx, n, ek, yout, dst0, s.W, s.nslots,
                     s.slot[0].mult, s.nslots > 1 ? s.slot[1].mult : 0, ng0, ng1, s.ngt_cap, s.G, s.GE);
}

__global__ void k_sub(double* a, const double* b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] -= b[i];
}

void launch_sub(double* a, const double* b, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_sub, dim3(cdivu(n, 256)), dim3(256), 0, st, a, b, n);
}

void launch_ej(const DevState& s, double* ej, hipStream_t st) {
  hipLaunchKernelGGL(k_ej, dim3(s.W), dim3(64), 0, st, s.fb, ej, s.G, s.GE);
}

void launch_predict(const DevState& s, const Mlp& m, const float* X, int n, double* S, hipStream_t st) {
  mlp_forward(m, X, nullptr, n, st);
  hipLaunchKernelGGL(k_pred_to_double, dim3(cdivu(n, 256)), dim3(256), 0, st, m.h[3], S, n, m.t_mu, m.t_sd);
  (void)s;
}

void launch_energy_cv(const DevState& s, const double* S, const double* ek, const float* pinv, double delta,
                      double soff, const double* center, double halfw, double* dres, hipStream_t st) {
  hipLaunchKernelGGL(k_energy_cv, dim3(s.W), dim3(64), 0, st, s.fb, s.e1, S, ek, pinv, s.weights, s.eloc, dres, s.G,
                     s.GE, s.nslots, s.slot[0].mult, s.nslots > 1 ? s.slot[1].mult : 0, s.ecore, delta, soff, center,
                     halfw);
}

}  // namespace afq
