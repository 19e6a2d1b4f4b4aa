// hermite4_wrap_g6.cpp -- shared-timestep Hermite-4 integrator written against
// the caller side of the GRAPE-6 header only: the by-value wrappers
// g6_open / g6_npipes / g6_set_ti / g6_set_j_particle / g6calc_firsthalf /
// g6calc_lasthalf / g6_close, the position-only g6_set_j_particle_multisend_mxfast_
// and the board-management stubs, in the call pattern of the reference's own
// integrator (testCodes/src/test_integrator.cpp:94-96, 218, 238, 267).
//
// usage: hermite4_wrap_g6 <input.bin> <t_end> <eta> <eps2>
//   input.bin: int32 n, then n x (mass, x, y, z, vx, vy, vz) doubles.
// prints "stubs <sum of stub returns>", "mxfast acc_pot_identical <0|1> jerk_differs <0|1>",
// one line per step and finally "final dE/E0 <value>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "g6lib.h"

namespace {
struct Sys {
  int n = 0;
  std::vector<double> m, x, v, a, j, pot, h2;
  std::vector<int> id;
};

// Forces on all particles at xp, vp in chunks of the pipe count.
void forces(Sys& s, std::vector<double>& xp, std::vector<double>& vp, double eps2, int npipes) {
  for (int i0 = 0; i0 < s.n; i0 += npipes) {
    const int ni = std::min(npipes, s.n - i0);
    auto X = reinterpret_cast<double(*)[3]>(&xp[3 * i0]);
    auto V = reinterpret_cast<double(*)[3]>(&vp[3 * i0]);
    auto A = reinterpret_cast<double(*)[3]>(&s.a[3 * i0]);
    auto J = reinterpret_cast<double(*)[3]>(&s.j[3 * i0]);
    g6calc_firsthalf(0, s.n, ni, &s.id[i0], X, V, A, J, &s.pot[i0], eps2, &s.h2[i0]);
    g6calc_lasthalf(0, s.n, ni, &s.id[i0], X, V, eps2, &s.h2[i0], A, J, &s.pot[i0]);
  }
}

void send_all(Sys& s, double t) {
  double k18[3] = {0, 0, 0};
  for (int i = 0; i < s.n; ++i) {
    double a2[3] = {0.5 * s.a[3 * i], 0.5 * s.a[3 * i + 1], 0.5 * s.a[3 * i + 2]};
    double j6[3] = {s.j[3 * i] / 6, s.j[3 * i + 1] / 6, s.j[3 * i + 2] / 6};
    g6_set_j_particle(0, i, s.id[i], t, t, s.m[i], k18, j6, a2, &s.v[3 * i], &s.x[3 * i]);
  }
}

double energy(const Sys& s) {
  double ek = 0, ep = 0;
  for (int i = 0; i < s.n; ++i) {
    const double* v = &s.v[3 * i];
    ek += 0.5 * s.m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    ep += 0.5 * s.m[i] * s.pot[i];
  }
  return ek + ep;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s input.bin t_end eta eps2\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  Sys s;
  if (fread(&s.n, sizeof(int), 1, f) != 1) return 2;
  const int n = s.n;
  s.m.resize(n); s.x.resize(3 * n); s.v.resize(3 * n); s.a.assign(3 * n, 0); s.j.assign(3 * n, 0);
  s.pot.assign(n, 0); s.h2.assign(n, 0); s.id.resize(n);
  for (int i = 0; i < n; ++i) {
    double r[7];
    if (fread(r, sizeof(double), 7, f) != 7) return 2;
    s.m[i] = r[0];
    for (int d = 0; d < 3; ++d) { s.x[3 * i + d] = r[1 + d]; s.v[3 * i + d] = r[4 + d]; }
    s.id[i] = 3000 + i;
  }
  fclose(f);
  const double t_end = atof(argv[2]), eta = atof(argv[3]), eps2 = atof(argv[4]);

  if (g6_open(0) != 0) { fprintf(stderr, "g6_open failed\n"); return 1; }
  g6_set_tunit(1.0);
  g6_set_xunit(1.0);
  const int npipes = g6_npipes();
  int stubs = g6_reset(0) + g6_reset_fofpga(0) + g6_reinitialize(0) + g6_print_chip_status(0) +
              g6_initialize_jp_buffer(0, n) + g6_flush_jp_buffer_and_multisend(0, 0);
  int cl = 0;
  stubs += g6_set_overflow_flag_test_mode_(&cl);
  printf("stubs %d npipes %d\n", stubs, npipes);

  // position-only load: acc and pot depend on positions and masses only, so
  // they must equal those of the full load bit for bit; jerk sees vj = 0
  int ncl = 1;
  for (int i = 0; i < n; ++i) g6_set_j_particle_multisend_mxfast_(&cl, &ncl, &i, &s.id[i], &s.m[i], &s.x[3 * i]);
  g6_set_ti(0, 0.0);
  forces(s, s.x, s.v, eps2, npipes);
  const std::vector<double> a_fast(s.a), j_fast(s.j), p_fast(s.pot);

  double t = 0.0;
  send_all(s, t);
  g6_set_ti(0, t);
  forces(s, s.x, s.v, eps2, npipes);
  const bool same = a_fast == s.a && p_fast == s.pot;
  const bool jdiff = j_fast != s.j;
  printf("mxfast acc_pot_identical %d jerk_differs %d\n", same ? 1 : 0, jdiff ? 1 : 0);

  const double e0 = energy(s);
  printf("start n %d E0 %.15e\n", n, e0);
  int step = 0;
  while (t < t_end) {
    double dt = 1e30;
    for (int i = 0; i < n; ++i) {
      const double* a = &s.a[3 * i];
      const double* jj = &s.j[3 * i];
      const double am = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      const double jm = std::sqrt(jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2]);
      if (jm > 0) dt = std::min(dt, eta * am / jm);
    }
    dt = std::pow(2.0, std::floor(std::log2(dt)));
    std::vector<double> xp(3 * n), vp(3 * n), a0(s.a), j0(s.j);
    for (int k = 0; k < 3 * n; ++k) {
      xp[k] = s.x[k] + dt * (s.v[k] + dt / 2 * (s.a[k] + dt / 3 * s.j[k]));
      vp[k] = s.v[k] + dt * (s.a[k] + dt / 2 * s.j[k]);
    }
    const double tn = t + dt;
    g6_set_ti(0, tn);
    forces(s, xp, vp, eps2, npipes);
    for (int k = 0; k < 3 * n; ++k) {
      const double vn = s.v[k] + dt / 2 * (a0[k] + s.a[k]) + dt * dt / 12 * (j0[k] - s.j[k]);
      s.x[k] = s.x[k] + dt / 2 * (s.v[k] + vn) + dt * dt / 12 * (a0[k] - s.a[k]);
      s.v[k] = vn;
    }
    t = tn;
    send_all(s, t);
    ++step;
    g6_set_ti(0, t);
    std::vector<double> ac(s.a), jc(s.j);
    forces(s, s.x, s.v, eps2, npipes);
    s.a = ac; s.j = jc;
    const double e = energy(s);
    printf("step %d t %.6f dt %.3e E %.15e dE/E0 %.3e\n", step, t, dt, e, (e - e0) / std::fabs(e0));
  }
  printf("final dE/E0 %.6e\n", (energy(s) - e0) / std::fabs(e0));
  g6_close(0);
  return 0;
}
