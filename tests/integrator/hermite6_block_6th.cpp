// hermite6_block_6th.cpp -- block-timestep Hermite-6 integrator (Nitadori &
// Makino 2008, New Astron. 13, 498) driving the 6th-order C-ABI
// (include/sapporo6thlib.h) as external 6th-order codes do: individual
// power-of-two steps; at each block time predict_all, the active particles'
// predicted state read back with pick_up_predictor_2, calc_force_on_predictors
// (acc, jerk, snap, pot) on them, the 6th-order corrector, crackle from the
// quintic through (a, j, s) at both ends of the step, set_j_particle for the
// corrected ones.  Only the 6th-order symbols are used.
//
// usage: hermite6_block_6th <input.bin> <t_end> <eta> <eps2>
//   input.bin: int32 n, then n x (mass, x, y, z, vx, vy, vz) doubles.
//   t_end is rounded up to a multiple of the largest step (1/16).
// prints "blocks <count> active_mean <mean> active_min <min>" and
// "final dE/E0 <value>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sapporo6thlib.h"

namespace {
constexpr double DT_MAX = 1.0 / 16;
constexpr double DT_MIN = 1.0 / (1 << 22);
constexpr long MAX_BLOCKS = 100000;

struct Sys {
  int n = 0;
  double eps2 = 0;
  std::vector<double> m, x, v, a, j, s, c, pot, t, dt;
  std::vector<int> id;
};

double norm3(const double* p) { return std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

void send(Sys& s, int i) {
  set_j_particle(i, &s.x[3 * i], &s.v[3 * i], &s.a[3 * i], &s.j[3 * i], &s.s[3 * i], &s.c[3 * i], s.m[i], s.t[i],
                 s.id[i], s.eps2);
}

// Forces on the listed particles at (xp, vp, ap) against addresses [0, n).
void forces(Sys& s, const std::vector<int>& act, std::vector<double>& xp, std::vector<double>& vp,
            std::vector<double>& ap, std::vector<double>& a, std::vector<double>& j, std::vector<double>& sn,
            std::vector<double>& pot) {
  const int na = (int)act.size();
  std::vector<int> ids(na), nnb(na);
  std::vector<double> mass(na), eps2(na, s.eps2), r2(na), crk(3 * na);
  for (int k = 0; k < na; ++k) { ids[k] = s.id[act[k]]; mass[k] = s.m[act[k]]; }
  a.assign(3 * na, 0.0); j.assign(3 * na, 0.0); sn.assign(3 * na, 0.0); pot.assign(na, 0.0);
  auto A3 = [](std::vector<double>& q) { return reinterpret_cast<double(*)[3]>(q.data()); };
  calc_force_on_predictors(na, s.n, ids.data(), A3(xp), A3(vp), A3(ap), mass.data(), eps2.data(), A3(a), A3(j),
                           A3(sn), A3(crk), pot.data(), nnb.data(), r2.data());
}

// Predicted state of the listed particles as stored on the device.
void pick_up(const std::vector<int>& act, std::vector<double>& xp, std::vector<double>& vp,
             std::vector<double>& ap) {
  const int na = (int)act.size();
  xp.resize(3 * na); vp.resize(3 * na); ap.resize(3 * na);
  for (int k = 0; k < na; ++k) {
    int id; double mass, eps2;
    pick_up_predictor_2(act[k], id, mass, eps2, &xp[3 * k], &vp[3 * k], &ap[3 * k]);
  }
}

// Largest power-of-two step <= the criterion, <= 2 dt_old, commensurate with t.
double quantise(double crit, double dt_old, double t) {
  double dt = DT_MAX;
  while (dt > crit && dt > DT_MIN) dt *= 0.5;
  if (dt_old > 0) dt = std::min(dt, 2 * dt_old);
  while (dt > DT_MIN && std::fmod(t, dt) != 0.0) dt *= 0.5;
  return dt;
}

// Aarseth-type criterion on acc, jerk, snap, crackle.
double criterion(const Sys& s, int i, double eta) {
  const double A = norm3(&s.a[3 * i]), J = norm3(&s.j[3 * i]), S = norm3(&s.s[3 * i]), C = norm3(&s.c[3 * i]);
  const double den = J * C + S * S;
  return den > 0 ? eta * std::sqrt((A * S + J * J) / den) : DT_MAX;
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
  s.m.resize(n); s.x.resize(3 * n); s.v.resize(3 * n);
  for (auto* q : {&s.a, &s.j, &s.s, &s.c}) q->assign(3 * n, 0.0);
  s.pot.assign(n, 0); s.id.resize(n); s.t.assign(n, 0.0); s.dt.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double r[7];
    if (fread(r, sizeof(double), 7, f) != 7) return 2;
    s.m[i] = r[0];
    for (int d = 0; d < 3; ++d) { s.x[3 * i + d] = r[1 + d]; s.v[3 * i + d] = r[4 + d]; }
    s.id[i] = i;
  }
  fclose(f);
  const double t_end = std::ceil(atof(argv[2]) / DT_MAX) * DT_MAX;
  const double eta = atof(argv[3]);
  s.eps2 = atof(argv[4]);
  initialize();

  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  std::vector<double> xp, vp, ap, a, jk, sn, pot;
  // two start-up evaluations: the snap needs the accelerations of both sides
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < n; ++i) send(s, i);
    predict_all(0.0, n);
    pick_up(all, xp, vp, ap);
    forces(s, all, xp, vp, ap, a, jk, sn, pot);
    s.a = a; s.j = jk; s.s = sn; s.pot = pot;
  }
  for (int i = 0; i < n; ++i) {
    const double A = norm3(&s.a[3 * i]), J = norm3(&s.j[3 * i]);
    s.dt[i] = quantise(J > 0 ? 0.01 * A / J : DT_MAX, 0.0, 0.0);
    send(s, i);
  }
  const double e0 = energy(s, s.pot);
  printf("start n %d E0 %.15e t_end %.6f\n", n, e0, t_end);

  long blocks = 0, active_sum = 0;
  int active_min = n;
  double t = 0.0;
  std::vector<int> act;
  while (t < t_end) {
    double tn = 1e30;
    for (int i = 0; i < n; ++i) tn = std::min(tn, s.t[i] + s.dt[i]);
    act.clear();
    for (int i = 0; i < n; ++i)
      if (s.t[i] + s.dt[i] == tn) act.push_back(i);
    const int na = (int)act.size();
    predict_all(tn, n);
    pick_up(act, xp, vp, ap);
    forces(s, act, xp, vp, ap, a, jk, sn, pot);
    for (int k = 0; k < na; ++k) {
      const int i = act[k];
      const double h = tn - s.t[i];
      for (int d = 0; d < 3; ++d) {
        const int q = 3 * i + d;
        const double a0 = s.a[q], j0 = s.j[q], s0 = s.s[q];
        const double a1 = a[3 * k + d], j1 = jk[3 * k + d], s1 = sn[3 * k + d];
        const double vn = s.v[q] + h / 2 * (a1 + a0) - h * h / 10 * (j1 - j0) + h * h * h / 120 * (s1 + s0);
        s.x[q] = s.x[q] + h / 2 * (vn + s.v[q]) - h * h / 10 * (a1 - a0) + h * h * h / 120 * (j1 + j0);
        s.v[q] = vn;
        // quintic through (a, j, s) at both ends: crackle, pop and the next derivative at the start
        const double R1 = a1 - a0 - j0 * h - s0 * h * h / 2, R2 = (j1 - j0 - s0 * h) * h, R3 = (s1 - s0) * h * h / 2;
        const double C = 10 * R1 - 4 * R2 + R3, P = -15 * R1 + 7 * R2 - 2 * R3, D = 6 * R1 - 3 * R2 + R3;
        const double c0 = 6 * C / (h * h * h), p0 = 24 * P / (h * h * h * h), d0 = 120 * D / (h * h * h * h * h);
        s.c[q] = c0 + p0 * h + d0 * h * h / 2;
        s.a[q] = a1; s.j[q] = j1; s.s[q] = s1;
      }
      s.pot[i] = pot[k];
      s.t[i] = tn;
      s.dt[i] = quantise(criterion(s, i, eta), s.dt[i], tn);
      if (tn + s.dt[i] > t_end) s.dt[i] = t_end - tn;   // land on t_end (a multiple of every step)
      send(s, i);
    }
    t = tn;
    ++blocks;
    active_sum += na;
    active_min = std::min(active_min, na);
    if (blocks % 1000 == 0) {
      double dmin = DT_MAX;
      for (int i = 0; i < n; ++i) dmin = std::min(dmin, s.dt[i]);
      printf("progress blocks %ld t %.6f dt_min %.3e\n", blocks, t, dmin);
      fflush(stdout);
    }
    if (blocks >= MAX_BLOCKS) {
      fprintf(stderr, "more than %ld block steps before t_end: step sizes collapsed\n", MAX_BLOCKS);
      return 1;
    }
  }
  // every particle is at t_end: potential of the corrected positions
  predict_all(t, n);
  pick_up(all, xp, vp, ap);
  forces(s, all, xp, vp, ap, a, jk, sn, pot);
  printf("blocks %ld active_mean %.1f active_min %d\n", blocks, (double)active_sum / std::max(1L, blocks),
         active_min);
  printf("final dE/E0 %.6e\n", (energy(s, pot) - e0) / std::fabs(e0));
  return 0;
}
