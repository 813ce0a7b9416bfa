// Drives the population through the C++ mirror classes (include/vmc_afqmc.hpp)
// the way calcMixedEstimatorLongProp drives DQMCWalker objects, on a system read
// from a binary file, and writes the final state; tests/test_gpu_cpp_api.py
// compares it with the CPU oracle.
//
// input  (little-endian): int32 N, na, nb, G, W, steps, walker_rhf, trial_rhf;
//        float64 ecore, dt, eshift; h1[N*N], h1mod[N*N], chol[G*N*N], phi_up[N*na],
//        phi_dn[N*nb], ref_up[N*na], ref_dn[N*nb] (prepProp's reference determinant),
//        dets[W*N*(na+nb)] complex, draws[steps*W*G]
// output: dets, overlaps (complex W), weights (W), hamAndOverlap (2 complex per W),
//         forceBias (W*G complex), localEnergy (W)
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>

#include "vmc_afqmc.hpp"

using vmc_afqmc::cplx;

template <class T>
static std::vector<T> rd(std::ifstream& f, size_t n) {
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
  if (!f) throw std::runtime_error("short input file");
  return v;
}

template <class T>
static void wr(std::ofstream& f, const std::vector<T>& v) {
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: api_parity input.bin output.bin\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    auto hd = rd<int32_t>(in, 8);
    const int N = hd[0], na = hd[1], nb = hd[2], G = hd[3], W = hd[4], steps = hd[5];
    const bool wrhf = hd[6] != 0, trhf = hd[7] != 0;
    auto sc = rd<double>(in, 3);
    auto h1 = rd<double>(in, (size_t)N * N);
    auto h1m = rd<double>(in, (size_t)N * N);
    auto chol = rd<double>(in, (size_t)G * N * N);
    auto up = rd<double>(in, (size_t)N * na);
    auto dn = rd<double>(in, (size_t)N * nb);
    auto refUp = rd<double>(in, (size_t)N * na);
    auto refDn = rd<double>(in, (size_t)N * nb);
    auto dets = rd<cplx>(in, (size_t)W * N * (na + nb));
    auto draws = rd<double>(in, (size_t)steps * W * G);

    vmc_afqmc::Hamiltonian ham(N, na, nb, G, sc[0], h1, h1m, chol);
    std::unique_ptr<vmc_afqmc::Wavefunction> wave;
    if (trhf) wave.reset(new vmc_afqmc::RHF(ham, up));
    else wave.reset(new vmc_afqmc::UHF(ham, up, dn));
    vmc_afqmc::DQMCWalkers walkers(ham, *wave, W, wrhf);
    walkers.prepProp(refUp, refDn, sc[1], 0.0);
    walkers.setDet(dets);
    for (int s = 0; s < steps; ++s) {
      std::vector<double> x(draws.begin() + (size_t)s * W * G, draws.begin() + (size_t)(s + 1) * W * G);
      walkers.propagatePhaseless(x, sc[2]);
      if (s == 1) walkers.orthogonalize();
    }
    std::vector<cplx> d, o;
    std::vector<double> w;
    walkers.getDet(d, o, w);
    auto ho = walkers.hamAndOverlap();
    auto fb = walkers.forceBias();
    auto el = walkers.localEnergy();
    std::ofstream out(argv[2], std::ios::binary);
    wr(out, d);
    wr(out, o);
    wr(out, w);
    wr(out, ho);
    wr(out, fb);
    wr(out, el);
    std::cout << "api_parity: " << W << " walkers, " << steps << " steps done\n";
  } catch (const vmc_afqmc::Error& e) {
    std::cerr << "vmc_afqmc error " << e.code << ": " << e.what() << "\n";
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
