// sixthlib.cpp -- 6th-order C-ABI (declared in include/sapporo6thlib.h) over
// one global engine, as the reference's sapporo6thlib.cpp (line 3).
#include "../../include/sapporo6thlib.h"

#include <algorithm>
#include <vector>

#include "devlist.h"
#include "engine.h"

namespace {
sap::Engine g_grav6;
std::vector<double> g_h2;   // zero neighbour radii (sapporo6thlib.cpp:34-39)
bool g_init = false;
}  // namespace

extern "C" {

void initialize_special(int ndev, int* list) {
  if (g_init) return;  // reference: only once (sapporo6thlib.cpp:21-24)
  const int prec = sap::precision_from_env(sapk::PREC_DOUBLE);
  if (g_grav6.open(sap::device_from_list(ndev, list), sapk::ORDER_H6, prec) == 0) g_init = true;
  g_h2.assign(sap::NPIPES, 0.0);
}

void initialize(void) {
  const int dev = sap::device_from_config_file();
  initialize_special(1, const_cast<int*>(&dev));
}

void set_j_particle(int add, double pos[3], double vel[3], double acc[3], double jrk[3], double snp[3],
                    double crk[3], double mass, double time, int id, double eps2) {
  g_grav6.set_j_particle(add, id, time, 0.0, mass, acc, jrk, vel, pos, snp, crk, eps2);
}

void predict_all(double time, int nj) {
  g_grav6.set_time(time);
  g_grav6.force_prediction(nj);
}

void no_predict_all(double time, int nj) {
  (void)time;
  g_grav6.set_no_time();
  g_grav6.force_prediction(nj);
}

void pick_up_predictor_2(int addr, int& id, double& mass, double& eps2, double pos[3], double vel[3],
                         double acc[3]) {
  double idt = 0.0;
  g_grav6.retrieve_predicted_j_particle(addr, &mass, &idt, &eps2, pos, vel, acc);
  id = (int)idt;
}

void calc_force_on_predictors(int ni, int nj, int ids[], double pos[][3], double vel[][3], double acc[][3],
                              double mass[], double eps2[], double accNew[][3], double jrkNew[][3],
                              double snpNew[][3], double crkNew[][3], double phi[], int nnb[],
                              double nnb_r2[]) {
  (void)mass;
  const int np = g_grav6.n_pipes();
  for (int j = 0; j < ni; j += np) {
    const int tni = std::min(np, ni - j);
    const double e2 = eps2[0];  // global softening (sapporo6thlib.cpp:182)
    g_grav6.start_grav_calc(nj, tni, &ids[j], &pos[j], &vel[j], &acc[j], e2, g_h2.data());
    g_grav6.get_grav_results(nj, tni, &accNew[j], &jrkNew[j], &snpNew[j], &crkNew[j], &phi[j], &nnb[j],
                             &nnb_r2[j], true);
  }
}

}  // extern "C"
