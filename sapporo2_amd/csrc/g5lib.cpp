// g5lib.cpp -- GRAPE-5 C-ABI (declared in include/g5lib.h) over one global
// engine, as the reference's sapporoG5lib.cpp does with `sapporo grav` (line 3).
//
// The reference opens the GRAPE5 order in FLOAT precision (sapporoG5lib.cpp:42):
// positions are used as floats.  Here positions are rounded to float on entry,
// so the double-single records carry them exactly (low parts zero) and the
// relative position is the float difference; masses and softening are float
// in the mixed kernel as well.  i-particles carry no id, so no pair is
// excluded and a particle's own mass contributes -m/eps to its potential
// (kernels.cu:565 skips the id test for GRAPE5).
#include "../../include/g5lib.h"

#include <climits>
#include <cstdio>
#include <vector>

#include "devlist.h"
#include "engine.h"

namespace {
sap::Engine g5_grav;
int g5_nj = -1;        // reference nj_to_use
double g5_eps2 = -1;   // reference eps_to_use
std::vector<double> g5_xi;

inline double to_float(double x) { return (double)(float)x; }

// i positions rounded to float, in a buffer the engine copies before returning
const double (*float_xi(int ni, const double (*xi)[3]))[3] {
  g5_xi.resize(3 * (size_t)ni);
  for (int i = 0; i < ni; ++i)
    for (int d = 0; d < 3; ++d) g5_xi[3 * (size_t)i + d] = to_float(xi[i][d]);
  return reinterpret_cast<const double(*)[3]>(g5_xi.data());
}

bool ready(const char* who) {
  if (g5_eps2 < 0) {
    fprintf(stderr, "sapporo2_amd: %s: eps not set with g5_set_eps_to_all; no force computed\n", who);
    return false;
  }
  if (g5_nj <= 0) {
    fprintf(stderr, "sapporo2_amd: %s: nj not set with g5_set_n; no force computed\n", who);
    return false;
  }
  return true;
}
}  // namespace

extern "C" {

int g5_open_special(int ndev, int* list) {
  g5_nj = -1;
  g5_eps2 = -1;
  return g5_grav.open(sap::device_from_list(ndev, list), sapk::ORDER_G5, sapk::PREC_DS);
}

void g5_open(void) {
  const int dev = sap::device_from_config_file();
  g5_open_special(1, const_cast<int*>(&dev));
}

void g5_close(void) { g5_grav.close(); }
int g5_get_number_of_pipelines(void) { return g5_grav.n_pipes(); }
void g5_set_n(int nj) { g5_nj = nj; }
void g5_set_eps2_to_all(double eps2) { g5_eps2 = eps2; }
void g5_set_eps_to_all(double eps) { g5_eps2 = eps * eps; }

// The reference indexes the caller's arrays by absolute address (sapporoG5lib.cpp:111-117).
void g5_set_xmj(int addr, int nj, double xj[][3], double jmass[]) {
  for (int i = addr; i < addr + nj; ++i) {
    const double x[3] = {to_float(xj[i][0]), to_float(xj[i][1]), to_float(xj[i][2])};
    g5_grav.set_j_particle(i, i, 0.0, 0.0, jmass[i], nullptr, nullptr, nullptr, x, nullptr, nullptr, 0.0);
  }
}

void g5_set_jp(int addr, int nj, double jmass[], double xj[][3]) { g5_set_xmj(addr, nj, xj, jmass); }

void g5_set_xi(int ni, double xi[][3]) {
  if (!ready("g5_set_xi")) return;
  if (ni > g5_grav.n_pipes()) {
    fprintf(stderr, "sapporo2_amd: g5_set_xi: ni = %d exceeds %d pipelines; no force computed\n", ni,
            g5_grav.n_pipes());
    return;
  }
  g5_grav.start_grav_calc(g5_nj, ni, nullptr, float_xi(ni, xi), nullptr, nullptr, g5_eps2, nullptr);
}

void g5_run(void) {}

void g5_get_force(int ni, double a[][3], double* phi) {
  g5_grav.get_grav_results(g5_nj, ni, a, nullptr, nullptr, nullptr, phi, nullptr, nullptr, false);
}

void g5_calculate_force_on_x(double xi[][3], double a[][3], double* phi, int ni) {
  if (!ready("g5_calculate_force_on_x")) return;
  const int np = g5_grav.n_pipes();
  for (int j = 0; j < ni; j += np) {
    const int n = ni - j < np ? ni - j : np;
    if (g5_grav.start_grav_calc(g5_nj, n, nullptr, float_xi(n, xi + j), nullptr, nullptr, g5_eps2, nullptr) != 0)
      return;
    g5_grav.get_grav_results(g5_nj, n, a + j, nullptr, nullptr, nullptr, phi + j, nullptr, nullptr, false);
  }
}

void g5_set_range(double xmin, double xmax, double mmin) { (void)xmin; (void)xmax; (void)mmin; }
int g5_get_jmemsize(void) { return INT_MAX; }
int g5_get_number_of_real_pipelines(void) { return 1; }
int g5_get_pcibus_freq(void) { return 1; }

}  // extern "C"
