// hermite4_g6.cpp -- shared-timestep Hermite-4 integrator driving the GRAPE-6
// C-ABI (include/g6lib.h), the way external GRAPE-6 codes and the reference's
// testCodes/src/test_integrator.cpp:76-182 do: set_j_particle for all, set_ti,
// firsthalf/lasthalf in chunks of g6_npipes, predictor-evaluator-corrector.
// It uses only g6_* symbols, so it links against any GRAPE-6 library unchanged.
//
// usage: hermite4_g6 <input.bin> <t_end> <eta> <eps2>
//   input.bin: int32 n, then n x (mass, x, y, z, vx, vy, vz) doubles.
// prints one line per step "step t dt E dE/E0" and finally "final dE/E0 <value>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "g6lib.h"

namespace {
struct Sys {
  int n = 0;
  std::vector<double> m, x, v, a, j, pot, h2;
  std::vector<int> id;
};

void forces(Sys& s, const std::vector<double>& xp, const std::vector<double>& vp, double eps2) {
  int cl = 0, nj = s.n;
  const int np = g6_npipes_();
  for (int i0 = 0; i0 < s.n; i0 += np) {
    int ni = std::min(np, s.n - i0);
    auto X = reinterpret_cast<double(*)[3]>(const_cast<double*>(&xp[3 * i0]));
    auto V = reinterpret_cast<double(*)[3]>(const_cast<double*>(&vp[3 * i0]));
    auto A = reinterpret_cast<double(*)[3]>(&s.a[3 * i0]);
    auto J = reinterpret_cast<double(*)[3]>(&s.j[3 * i0]);
    g6calc_firsthalf_(&cl, &nj, &ni, &s.id[i0], X, V, A, J, &s.pot[i0], &eps2, &s.h2[i0]);
    g6calc_lasthalf_(&cl, &nj, &ni, &s.id[i0], X, V, &eps2, &s.h2[i0], A, J, &s.pot[i0]);
  }
}

void send_all(Sys& s, double t) {
  int cl = 0;
  double zero3[3] = {0, 0, 0};
  for (int i = 0; i < s.n; ++i) {
    double a2[3] = {0.5 * s.a[3 * i], 0.5 * s.a[3 * i + 1], 0.5 * s.a[3 * i + 2]};
    double j6[3] = {s.j[3 * i] / 6, s.j[3 * i + 1] / 6, s.j[3 * i + 2] / 6};
    g6_set_j_particle_(&cl, &i, &s.id[i], &t, &t, &s.m[i], zero3, j6, a2, &s.v[3 * i], &s.x[3 * i]);
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
    s.id[i] = i;
  }
  fclose(f);
  const double t_end = atof(argv[2]), eta = atof(argv[3]), eps2 = atof(argv[4]);
  int dev = 0;
  if (g6_open_(&dev) != 0) { fprintf(stderr, "g6_open_ failed\n"); return 1; }
  double t = 0.0;
  send_all(s, t);
  g6_set_ti_(&dev, &t);
  forces(s, s.x, s.v, eps2);
  const double e0 = energy(s);
  printf("start n %d E0 %.15e\n", n, e0);
  int step = 0;
  while (t < t_end) {
    // shared timestep from the Aarseth criterion on |a|, |j|
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
    double tn = t + dt;
    g6_set_ti_(&dev, &tn);
    forces(s, xp, vp, eps2);
    // Hermite corrector
    for (int k = 0; k < 3 * n; ++k) {
      const double vn = s.v[k] + dt / 2 * (a0[k] + s.a[k]) + dt * dt / 12 * (j0[k] - s.j[k]);
      s.x[k] = s.x[k] + dt / 2 * (s.v[k] + vn) + dt * dt / 12 * (a0[k] - s.a[k]);
      s.v[k] = vn;
    }
    t = tn;
    send_all(s, t);
    ++step;
    // energy with the potential at the corrected positions (one extra evaluation)
    g6_set_ti_(&dev, &t);
    std::vector<double> ac(s.a), jc(s.j);
    forces(s, s.x, s.v, eps2);
    s.a = ac; s.j = jc;
    const double e = energy(s);
    printf("step %d t %.6f dt %.3e E %.15e dE/E0 %.3e\n", step, t, dt, e, (e - e0) / std::fabs(e0));
  }
  printf("final dE/E0 %.6e\n", (energy(s) - e0) / std::fabs(e0));
  g6_close_(&dev);
  return 0;
}
