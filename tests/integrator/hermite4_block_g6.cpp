// hermite4_block_g6.cpp -- block-timestep Hermite-4 integrator driving the
// GRAPE-6 C-ABI (include/g6lib.h) as external GRAPE-6 codes do: individual
// power-of-two steps, at each block time g6_set_ti (the library predicts all
// j-particles), firsthalf/lasthalf on the active particles in chunks of
// g6_npipes, Hermite corrector, g6_set_j_particle for the corrected ones.
// Only g6_* symbols are used, so it links against any GRAPE-6 library.
//
// usage: hermite4_block_g6 <input.bin> <t_end> <eta> <eps2>
//   input.bin: int32 n, then n x (mass, x, y, z, vx, vy, vz) doubles.
//   t_end is rounded up to a multiple of the largest step (1/16), where every
//   particle is synchronised.
// prints "blocks <count> active_mean <mean> active_min <min>" and
// "final dE/E0 <value>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "g6lib.h"

namespace {
constexpr double DT_MAX = 1.0 / 16;
constexpr double DT_MIN = 1.0 / (1 << 22);

struct Sys {
  int n = 0;
  std::vector<double> m, x, v, a, j, pot, h2, t, dt;
  std::vector<int> id;
};

double norm3(const double* p) { return std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// Forces on the listed particles at positions xp, vp (n_act x 3) against all j.
void forces(Sys& s, const std::vector<int>& act, const std::vector<double>& xp, const std::vector<double>& vp,
            std::vector<double>& a, std::vector<double>& j, std::vector<double>& pot, double eps2) {
  int cl = 0, nj = s.n;
  const int np = g6_npipes_();
  const int na = (int)act.size();
  std::vector<int> ids(na);
  for (int k = 0; k < na; ++k) ids[k] = s.id[act[k]];
  a.assign(3 * na, 0.0); j.assign(3 * na, 0.0); pot.assign(na, 0.0);
  std::vector<double> h2(na, 0.0);
  for (int i0 = 0; i0 < na; i0 += np) {
    int ni = std::min(np, na - i0);
    auto X = reinterpret_cast<double(*)[3]>(const_cast<double*>(&xp[3 * i0]));
    auto V = reinterpret_cast<double(*)[3]>(const_cast<double*>(&vp[3 * i0]));
    auto A = reinterpret_cast<double(*)[3]>(&a[3 * i0]);
    auto J = reinterpret_cast<double(*)[3]>(&j[3 * i0]);
    g6calc_firsthalf_(&cl, &nj, &ni, &ids[i0], X, V, A, J, &pot[i0], &eps2, &h2[i0]);
    g6calc_lasthalf_(&cl, &nj, &ni, &ids[i0], X, V, &eps2, &h2[i0], A, J, &pot[i0]);
  }
}

void send(Sys& s, int i) {
  int cl = 0;
  double zero3[3] = {0, 0, 0};
  double a2[3] = {0.5 * s.a[3 * i], 0.5 * s.a[3 * i + 1], 0.5 * s.a[3 * i + 2]};
  double j6[3] = {s.j[3 * i] / 6, s.j[3 * i + 1] / 6, s.j[3 * i + 2] / 6};
  g6_set_j_particle_(&cl, &i, &s.id[i], &s.t[i], &s.dt[i], &s.m[i], zero3, j6, a2, &s.v[3 * i], &s.x[3 * i]);
}

// Largest power-of-two step <= the criterion, <= 2 dt_old, commensurate with t.
double quantise(double crit, double dt_old, double t) {
  double dt = DT_MAX;
  while (dt > crit && dt > DT_MIN) dt *= 0.5;
  if (dt_old > 0) dt = std::min(dt, 2 * dt_old);
  while (dt > DT_MIN && std::fmod(t, dt) != 0.0) dt *= 0.5;
  return dt;
}

double energy(const Sys& s, const std::vector<double>& pot) {
  double ek = 0, ep = 0;
  for (int i = 0; i < s.n; ++i) {
    const double* v = &s.v[3 * i];
    ek += 0.5 * s.m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    ep += 0.5 * s.m[i] * pot[i];
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
  s.pot.assign(n, 0); s.h2.assign(n, 0); s.id.resize(n); s.t.assign(n, 0.0); s.dt.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double r[7];
    if (fread(r, sizeof(double), 7, f) != 7) return 2;
    s.m[i] = r[0];
    for (int d = 0; d < 3; ++d) { s.x[3 * i + d] = r[1 + d]; s.v[3 * i + d] = r[4 + d]; }
    s.id[i] = i;
  }
  fclose(f);
  const double t_end = std::ceil(atof(argv[2]) / DT_MAX) * DT_MAX;
  const double eta = atof(argv[3]), eps2 = atof(argv[4]);
  int dev = 0;
  if (g6_open_(&dev) != 0) { fprintf(stderr, "g6_open_ failed\n"); return 1; }

  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  for (int i = 0; i < n; ++i) send(s, i);
  double t0 = 0.0;
  g6_set_ti_(&dev, &t0);
  std::vector<double> a, jk, pot;
  forces(s, all, s.x, s.v, a, jk, pot, eps2);
  s.a = a; s.j = jk; s.pot = pot;
  for (int i = 0; i < n; ++i) {
    const double jm = norm3(&s.j[3 * i]);
    s.dt[i] = quantise(jm > 0 ? eta * norm3(&s.a[3 * i]) / jm : DT_MAX, 0.0, 0.0);
    send(s, i);
  }
  const double e0 = energy(s, s.pot);
  printf("start n %d E0 %.15e t_end %.6f\n", n, e0, t_end);

  long blocks = 0, active_sum = 0;
  int active_min = n;
  double t = 0.0;
  std::vector<int> act;
  std::vector<double> xp, vp;
  while (t < t_end) {
    double tn = 1e30;
    for (int i = 0; i < n; ++i) tn = std::min(tn, s.t[i] + s.dt[i]);
    act.clear();
    for (int i = 0; i < n; ++i)
      if (s.t[i] + s.dt[i] == tn) act.push_back(i);
    const int na = (int)act.size();
    xp.resize(3 * na); vp.resize(3 * na);
    for (int k = 0; k < na; ++k) {
      const int i = act[k];
      const double h = tn - s.t[i];
      for (int d = 0; d < 3; ++d) {
        const int q = 3 * i + d;
        xp[3 * k + d] = s.x[q] + h * (s.v[q] + h / 2 * (s.a[q] + h / 3 * s.j[q]));
        vp[3 * k + d] = s.v[q] + h * (s.a[q] + h / 2 * s.j[q]);
      }
    }
    g6_set_ti_(&dev, &tn);
    forces(s, act, xp, vp, a, jk, pot, eps2);
    for (int k = 0; k < na; ++k) {
      const int i = act[k];
      const double h = tn - s.t[i];
      for (int d = 0; d < 3; ++d) {
        const int q = 3 * i + d;
        const double a1 = a[3 * k + d], j1 = jk[3 * k + d];
        const double vn = s.v[q] + h / 2 * (s.a[q] + a1) + h * h / 12 * (s.j[q] - j1);
        s.x[q] = s.x[q] + h / 2 * (s.v[q] + vn) + h * h / 12 * (s.a[q] - a1);
        s.v[q] = vn;
        s.a[q] = a1;
        s.j[q] = j1;
      }
      s.pot[i] = pot[k];
      s.t[i] = tn;
      const double jm = norm3(&s.j[3 * i]);
      s.dt[i] = quantise(jm > 0 ? eta * norm3(&s.a[3 * i]) / jm : DT_MAX, s.dt[i], tn);
      if (tn + s.dt[i] > t_end) s.dt[i] = t_end - tn;   // land on t_end (a multiple of every step)
      send(s, i);
    }
    t = tn;
    ++blocks;
    active_sum += na;
    active_min = std::min(active_min, na);
  }
  // every particle is at t_end: potential of the corrected positions
  g6_set_ti_(&dev, &t);
  forces(s, all, s.x, s.v, a, jk, pot, eps2);
  printf("blocks %ld active_mean %.1f active_min %d\n", blocks, (double)active_sum / std::max(1L, blocks),
         active_min);
  printf("final dE/E0 %.6e\n", (energy(s, pot) - e0) / std::fabs(e0));
  g6_close_(&dev);
  return 0;
}
