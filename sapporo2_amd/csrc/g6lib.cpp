// g6lib.cpp -- GRAPE-6 C-ABI (declared in include/g6lib.h) over one global
// engine, as the reference's sapporoG6lib.cpp does with `sapporo grav` (line 3).
#include "../../include/g6lib.h"
#include "../../include/sapporo_amd.h"

#include <cstdio>
#include <vector>

#include "devlist.h"
#include "engine.h"

namespace {
sapamd_engine g_handle;
sap::Engine& g_grav = g_handle.eng;
std::vector<double> g_dsmin;  // reference dsmin_i scratch (sapporoG6lib.cpp:18, 41)
}  // namespace

extern "C" {

sapamd_engine* sapamd_g6_engine(void) { return &g_handle; }

int g6_open_special(int ndev, int* list) {
  const int prec = sap::precision_from_env(sapk::PREC_DS);
  const int res = g_grav.open(sap::device_from_list(ndev, list), sapk::ORDER_H4, prec);
  g_dsmin.assign(sap::NPIPES, 0.0);
  return res;
}

int g6_open_(int* id) {
  (void)id;
  const int dev = sap::device_from_config_file();
  return g6_open_special(1, const_cast<int*>(&dev));
}

int g6_close_(int* id) {
  (void)id;
  return g_grav.close();
}
int g6_npipes_(void) { return g_grav.n_pipes(); }
int g6_set_tunit_(double*) { return 0; }
int g6_set_xunit_(double*) { return 0; }
int g6_set_ti_(int* id, double* ti) {
  (void)id;
  return g_grav.set_time(*ti);
}

// The GRAPE-6 j input is a/2 and j/6 (sapporoG6lib.cpp:103-104).
int g6_set_j_particle_(int* cluster_id, int* address, int* index, double* tj, double* dtj, double* mass,
                       double k18[3], double j6[3], double a2[3], double v[3], double x[3]) {
  (void)cluster_id;
  (void)k18;
  const double acc[3] = {a2[0] * 2, a2[1] * 2, a2[2] * 2};
  const double jrk[3] = {j6[0] * 6, j6[1] * 6, j6[2] * 6};
  return g_grav.set_j_particle(*address, *index, *tj, *dtj, *mass, acc, jrk, v, x, nullptr, nullptr, 0.0);
}

void g6calc_firsthalf_(int* cluster_id, int* nj, int* ni, int index[], double xi[][3], double vi[][3],
                       double aold[][3], double j6old[][3], double phiold[3], double* eps2, double h2[]) {
  (void)cluster_id; (void)j6old; (void)phiold;
  g_grav.start_grav_calc(*nj, *ni, index, xi, vi, aold, *eps2, h2);
}

int g6calc_lasthalf_(int* cluster_id, int* nj, int* ni, int index[], double xi[][3], double vi[][3],
                     double* eps2, double h2[], double acc[][3], double jerk[][3], double pot[]) {
  (void)cluster_id; (void)index; (void)xi; (void)vi; (void)eps2; (void)h2;
  return g_grav.get_grav_results(*nj, *ni, acc, jerk, nullptr, nullptr, pot, nullptr, nullptr, false);
}

int g6calc_lasthalf2_(int* cluster_id, int* nj, int* ni, int index[], double xi[][3], double vi[][3],
                      double* eps2, double h2[], double acc[][3], double jerk[][3], double pot[], int* inn) {
  (void)cluster_id; (void)index; (void)xi; (void)vi; (void)eps2; (void)h2;
  if ((int)g_dsmin.size() < *ni) g_dsmin.resize(*ni);
  return g_grav.get_grav_results(*nj, *ni, acc, jerk, nullptr, nullptr, pot, inn, g_dsmin.data(), true);
}

void get_j_part_data(int addr, int nj, double* pos, double* vel, double* acc, double* jrk, double* ppos,
                     double* pvel, double& mass, double& eps2, int& id) {
  (void)nj;
  double idt = 0.0;
  g_grav.retrieve_j_particle_state(addr, &mass, &idt, &eps2, pos, vel, acc, jrk, ppos, pvel, nullptr);
  id = (int)idt;
}

int g6_initialize_jp_buffer_(int* cluster_id, int* buf_size) {
  (void)cluster_id; (void)buf_size;
  return 0;
}
int g6_flush_jp_buffer_(int* cluster_id) {
  (void)cluster_id;
  return 0;
}
int g6_reset_(int* cluster_id) {
  (void)cluster_id;
  return 0;
}
int g6_reset_fofpga_(int* cluster_id) {
  (void)cluster_id;
  return 0;
}

int g6_read_neighbour_list_(int* cluster_id) {
  (void)cluster_id;
  return g_grav.read_ngb_list();
}

int g6_get_neighbour_list_(int* cluster_id, int* ipipe, int* maxlength, int* n_neighbours,
                           int neighbour_list[]) {
  (void)cluster_id;
  return g_grav.get_ngb_list(*ipipe, *maxlength, n_neighbours, neighbour_list);
}

}  // extern "C"
