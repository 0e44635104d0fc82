// qd_fixture_run - command-line driver over the C ABI of libqdamd.
//
// Plays the role of the reference's Matlab driver plus MEX entry
// (FH2main.m:195-215 building FH2Data, cudaOpenMPEvolution.C:15-79 parsing it
// and running the propagation): it reads one job's inputs from a fixture
// file, fills a qd_inputs struct exactly as a MEX shim would from the
// mxArray fields, calls qd_evolve, and writes the in-place outputs.
//
//   qd_fixture_run <fixture> <output>
//
// Fixture (little endian), written by qd-openmp-cuda_amd/inputs.py
// write_fixture():
//   char[8]  "QDFX0001"
//   int32[17] n1 n2 ntheta m n_omegas J parity l_max legendre_n_max
//             total_steps steps k_max n_energies n_dividing_surface
//             n_gradient_points calculate_CRP has_dump
//   double[7] r1.left r1.dr r1.mass r2.left r2.dr r2.mass time_step
//   double x[ntheta], w[ntheta], pot[n1 n2 ntheta]
//   int32  omegas[n_omegas]
//   double associated_legendres[ntheta legendre_n_max n_omegas]
//   double wave_packets[2 n1 n2 ntheta n_omegas]   (interleaved complex)
//   double dump1[n1], dump2[n2]                     (when has_dump)
//   double energies[n_energies], eta_sq[n_energies]
// Output:
//   char[8] "QDOUT001", int32 steps,
//   double observables[total_steps n_omegas 5],
//   double wave_packets[2 n1 n2 ntheta n_omegas], double CRP[n_energies]
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "qd.h"

namespace {

struct Reader {
  FILE *f;
  bool ok = true;
  template <class T>
  void read(T *dst, size_t n) {
    if (n && std::fread(dst, sizeof(T), n, f) != n) ok = false;
  }
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    read(v.data(), n);
    return v;
  }
};

int usage() {
  std::fprintf(stderr, "usage: qd_fixture_run <fixture> <output>\n");
  return 2;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) return usage();
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  Reader r{f};
  char magic[8];
  r.read(magic, 8);
  if (!r.ok || std::memcmp(magic, "QDFX0001", 8) != 0) {
    std::fprintf(stderr, "%s is not a fixture file\n", argv[1]);
    return 1;
  }
  int32_t h[17];
  r.read(h, 17);
  double g[7];
  r.read(g, 7);
  const int n1 = h[0], n2 = h[1], nt = h[2], m = h[3], nO = h[4];
  const int lnmax = h[8], total = h[9], nE = h[12];
  const bool has_dump = h[16] != 0;
  if (!r.ok || n1 <= 0 || n2 <= 0 || nt <= 0 || nO <= 0 || lnmax <= 0 || total < 0 || nE < 0) {
    std::fprintf(stderr, "bad fixture header\n");
    return 1;
  }
  const size_t G = (size_t)n1 * n2 * nt;
  auto x = r.vec<double>(nt);
  auto w = r.vec<double>(nt);
  auto pot = r.vec<double>(G);
  auto omegas = r.vec<int32_t>(nO);
  auto P = r.vec<double>((size_t)nt * lnmax * nO);
  auto psi = r.vec<double>(2 * G * nO);
  std::vector<double> d1, d2;
  if (has_dump) {
    d1 = r.vec<double>(n1);
    d2 = r.vec<double>(n2);
  }
  auto energies = r.vec<double>(nE);
  auto eta = r.vec<double>(nE);
  std::fclose(f);
  if (!r.ok) {
    std::fprintf(stderr, "truncated fixture\n");
    return 1;
  }

  // the struct a MEX shim fills from FH2Data (matlabStructures.C:17-153)
  qd_inputs in;
  std::memset(&in, 0, sizeof in);
  int32_t steps = h[10];
  std::vector<double> obs((size_t)total * nO * QD_OBS_FIELDS);
  std::vector<double> crp(nE);
  in.r1 = qd_radial{n1, g[0], g[1], g[2]};
  in.r2 = qd_radial{n2, g[3], g[4], g[5]};
  in.theta.n = nt;
  in.theta.m = m;
  in.theta.x = x.data();
  in.theta.w = w.data();
  in.pot = pot.data();
  in.time.total_steps = total;
  in.time.time_step = g[6];
  in.time.steps = &steps;
  in.dump1.dump = has_dump ? d1.data() : nullptr;
  in.dump2.dump = has_dump ? d2.data() : nullptr;
  in.crp.energies = energies.data();
  in.crp.eta_sq = eta.data();
  in.crp.CRP = crp.data();
  in.crp.n_dividing_surface = h[13];
  in.crp.n_gradient_points = h[14];
  in.crp.n_energies = nE;
  in.crp.calculate_CRP = h[15];
  in.omega_states.J = h[5];
  in.omega_states.parity = h[6];
  in.omega_states.l_max = h[7];
  in.omega_states.n_omegas_max = nO;
  in.omega_states.n_omegas = nO;
  in.omega_states.omegas = omegas.data();
  in.omega_states.legendre_n_max = lnmax;
  in.omega_states.associated_legendres = P.data();
  in.omega_states.wave_packets = psi.data();
  in.k_max = h[11];
  in.precision = QD_PREC_FP64;
  in.final_copy = 1;
  in.observables = obs.data();

  const qd_status st = qd_evolve(&in);
  if (st != QD_OK) {
    std::fprintf(stderr, "qd_evolve failed (%d): %s\n", (int)st, qd_last_global_error());
    return 1;
  }
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) {
    std::fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  bool wok = std::fwrite("QDOUT001", 1, 8, o) == 8;
  wok = wok && std::fwrite(&steps, sizeof steps, 1, o) == 1;
  wok = wok && std::fwrite(obs.data(), sizeof(double), obs.size(), o) == obs.size();
  wok = wok && std::fwrite(psi.data(), sizeof(double), psi.size(), o) == psi.size();
  wok = wok && std::fwrite(crp.data(), sizeof(double), crp.size(), o) == crp.size();
  wok = (std::fclose(o) == 0) && wok;
  if (!wok) {
    std::fprintf(stderr, "short write to %s\n", argv[2]);
    return 1;
  }
  double norm = 0.0;
  for (int b = 0; b < nO && total > 0; b++) norm += obs[((size_t)(total - 1) * nO + b) * QD_OBS_FIELDS + QD_OBS_NORM];
  std::printf("%s: %d steps, time.steps = %d, final norm %.15f\n", qd_version(), total, steps, norm);
  return 0;
}
