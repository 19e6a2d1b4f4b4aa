// capi.cpp -- handle-based C-ABI over sap::Engine (declared in include/sapporo_amd.h).
#include "../../include/sapporo_amd.h"

#include <new>

#include "engine.h"

using V3 = const double (*)[3];
using W3 = double (*)[3];

extern "C" {

sapamd_engine* sapamd_create(void) { return new (std::nothrow) sapamd_engine(); }
void sapamd_destroy(sapamd_engine* e) { delete e; }

int sapamd_open(sapamd_engine* e, int device, int order, int precision) {
  return e ? e->eng.open(device, order, precision) : -1;
}
int sapamd_close(sapamd_engine* e) { return e ? e->eng.close() : -1; }
int sapamd_get_n_pipes(sapamd_engine* e) { return e ? e->eng.n_pipes() : sap::NPIPES; }
int sapamd_set_time(sapamd_engine* e, double t) { return e ? e->eng.set_time(t) : -1; }
int sapamd_set_no_time(sapamd_engine* e) { return e ? e->eng.set_no_time() : -1; }

int sapamd_set_j_particle(sapamd_engine* e, int address, int id, double tj, double dtj, double mass,
                          const double acc[3], const double jrk[3], const double vel[3],
                          const double pos[3], const double snp[3], const double crk[3], double eps) {
  return e ? e->eng.set_j_particle(address, id, tj, dtj, mass, acc, jrk, vel, pos, snp, crk, eps) : -1;
}

int sapamd_set_j_particles(sapamd_engine* e, int n, const int* address, const int* id, const double* tj,
                           const double* dtj, const double* mass, const double* acc, const double* jrk,
                           const double* vel, const double* pos, const double* snp, const double* crk,
                           const double* eps) {
  return e ? e->eng.set_j_particles(n, address, id, tj, dtj, mass, acc, jrk, vel, pos, snp, crk, eps) : -1;
}

int sapamd_start_grav_calc(sapamd_engine* e, int nj, int ni, const int* id, const double* xi,
                           const double* vi, const double* ai, double eps2, const double* h2) {
  return e ? e->eng.start_grav_calc(nj, ni, id, reinterpret_cast<V3>(xi), reinterpret_cast<V3>(vi),
                                    reinterpret_cast<V3>(ai), eps2, h2)
           : -1;
}

int sapamd_get_grav_results(sapamd_engine* e, int nj, int ni, double* acc, double* jrk, double* snp,
                            double* crk, double* pot, int* nnb, double* dsmin, int ngb) {
  return e ? e->eng.get_grav_results(nj, ni, reinterpret_cast<W3>(acc), reinterpret_cast<W3>(jrk),
                                     reinterpret_cast<W3>(snp), reinterpret_cast<W3>(crk), pot, nnb,
                                     dsmin, ngb != 0)
           : -1;
}

int sapamd_set_clusters(sapamd_engine* e, int n, const int* nj) { return e ? e->eng.set_clusters(n, nj) : -1; }
int sapamd_cluster_offset(sapamd_engine* e, int c) { return e ? e->eng.cluster_offset(c) : -1; }
int sapamd_start_grav_calc_clusters(sapamd_engine* e, int ni, const int* cluster, const int* id, const double* xi,
                                    const double* vi, const double* ai, const double* eps2, const double* h2) {
  return e ? e->eng.start_grav_calc_clusters(ni, cluster, id, reinterpret_cast<V3>(xi), reinterpret_cast<V3>(vi),
                                             reinterpret_cast<V3>(ai), eps2, h2)
           : -1;
}

int sapamd_read_ngb_list(sapamd_engine* e) { return e ? e->eng.read_ngb_list() : -1; }
int sapamd_get_ngb_list(sapamd_engine* e, int ipipe, int maxlen, int* n, int* list) {
  return e ? e->eng.get_ngb_list(ipipe, maxlen, n, list) : -1;
}
int sapamd_force_prediction(sapamd_engine* e, int nj) { return e ? e->eng.force_prediction(nj) : -1; }

int sapamd_retrieve_predicted_j_particle(sapamd_engine* e, int addr, double* mass, double* id,
                                         double* eps2, double* pos, double* vel, double* acc) {
  return e ? e->eng.retrieve_predicted_j_particle(addr, mass, id, eps2, pos, vel, acc) : -1;
}

int sapamd_retrieve_j_particle_state(sapamd_engine* e, int addr, double* mass, double* id, double* eps2,
                                     double* pos, double* vel, double* acc, double* jrk, double* ppos,
                                     double* pvel, double* pacc) {
  return e ? e->eng.retrieve_j_particle_state(addr, mass, id, eps2, pos, vel, acc, jrk, ppos, pvel, pacc)
           : -1;
}

int sapamd_set_force_override(sapamd_engine* e, int P, int S) {
  if (!e) return -1;
  e->eng.set_force_override(P, S);
  return 0;
}
int sapamd_set_split_major(sapamd_engine* e, int on) {
  if (!e) return -1;
  e->eng.set_split_major(on != 0);
  return 0;
}
int sapamd_set_ngb_pipeline(sapamd_engine* e, int on) {
  if (!e) return -1;
  e->eng.set_ngb_pipeline(on != 0);
  return 0;
}
int sapamd_set_timing(sapamd_engine* e, int on) {
  if (!e) return -1;
  e->eng.set_timing(on != 0);
  return 0;
}

int sapamd_retrieve_predicted_j_particles(sapamd_engine* e, int n, const int* addr, double* mass,
                                          double (*pos)[3], double (*vel)[3], double (*acc)[3]) {
  if (!e) return -1;
  return e->eng.retrieve_predicted_j_particles(n, addr, mass, pos, vel, acc);
}

long long sapamd_state_bytes(sapamd_engine* e) {
  if (!e) return -1;
  return e->eng.state_bytes();
}

int sapamd_save_state(sapamd_engine* e, void* buf, long long bytes) {
  if (!e) return -1;
  return e->eng.save_state(buf, bytes);
}

int sapamd_load_state(sapamd_engine* e, const void* buf, long long bytes) {
  if (!e) return -1;
  return e->eng.load_state(buf, bytes);
}

int sapamd_set_graphs(sapamd_engine* e, int on) {
  if (!e) return -1;
  e->eng.set_graphs(on != 0);
  return 0;
}

int sapamd_graph_count(sapamd_engine* e) {
  if (!e) return -1;
  return e->eng.graph_count();
}
float sapamd_last_force_ms(sapamd_engine* e) { return e ? e->eng.last_force_ms() : 0.f; }
float sapamd_last_predict_ms(sapamd_engine* e) { return e ? e->eng.last_predict_ms() : 0.f; }
float sapamd_last_scatter_ms(sapamd_engine* e, int* nupd) {
  if (!e) {
    if (nupd) *nupd = 0;
    return 0.f;
  }
  return e->eng.last_scatter_ms(nupd);
}
float sapamd_last_finalize_ms(sapamd_engine* e) { return e ? e->eng.last_finalize_ms() : 0.f; }
float sapamd_last_device_ms(sapamd_engine* e) { return e ? e->eng.last_device_ms() : 0.f; }
int sapamd_last_config(sapamd_engine* e, int* P, int* S, int* jchunk) {
  if (!e) return -1;
  const sapk::ForceConfig c = e->eng.last_config();
  if (P) *P = c.P;
  if (S) *S = c.S;
  if (jchunk) *jchunk = c.jchunk;
  return 0;
}
int sapamd_load_tuning(sapamd_engine* e, const char* path) { return e ? e->eng.load_tuning(path) : -1; }
int sapamd_tuning_entries(sapamd_engine* e) { return e ? e->eng.tuning_entries() : 0; }
int sapamd_num_cus(sapamd_engine* e) { return e ? e->eng.num_cus() : 0; }
int sapamd_synchronize(sapamd_engine* e) { return e ? e->eng.synchronize() : -1; }
const char* sapamd_last_error(sapamd_engine* e) { return e ? e->eng.last_error() : "null handle"; }

}  // extern "C"
