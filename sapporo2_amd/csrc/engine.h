// engine.h -- host engine of the direct-summation library.
//
// One Engine drives one GPU.  It mirrors the reference's host class
// (lib/include/sapporohostclass.h:121-176, lib/src/sapporohostclass.cpp) with
// the same call sequence and result semantics, re-designed for one MI355X:
//   * j-particles live in an FP64 field-major store in HBM; updates are staged
//     in pinned host memory and sent with one asynchronous copy per flush,
//     then scattered by a kernel (no per-particle map lookups on the device);
//   * each force call is scatter -> [predict] -> force -> finalize on one HIP
//     stream; the finalize writes the result rows and the sorted, packed
//     neighbour lists straight into host-mapped pinned memory, so lasthalf and
//     read_neighbour_list wait for one completion event and copy nothing.  The
//     kernels take the call's values (block time, softening, counts) from a
//     small device-side descriptor, so with graphs on the sequence is captured
//     once per launch shape and replayed (one graph launch per call).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "kernels.h"

namespace sap {

constexpr int NPIPES = 16384;        // reference defines.h:13-15
constexpr int NJ_INITIAL = 16384;    // reference sapporohostclass.cpp:161-162
constexpr int ENS_MAX_I = 1 << 22;   // i-particles per ensemble call

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // order: sapk::ORDER_H4 / ORDER_H6; precision: PREC_DS (mixed) or PREC_DOUBLE.
  int open(int device, int order, int precision);
  int close();
  bool is_open() const { return open_; }
  int order() const { return order_; }
  int precision() const { return f64_ ? sapk::PREC_DOUBLE : sapk::PREC_DS; }
  int n_pipes() const { return NPIPES; }

  int set_time(double t);
  int set_no_time();
  int set_j_particle(int addr, int id, double tj, double dtj, double mass, const double a[3],
                     const double jrk[3], const double v[3], const double x[3], const double snp[3],
                     const double crk[3], double eps);
  // Array form of set_j_particle for n particles (vectors are n x 3 row-major;
  // tj, dtj, eps, snp, crk may be NULL).  Same per-particle semantics.
  int set_j_particles(int n, const int* addr, const int* id, const double* tj, const double* dtj,
                      const double* mass, const double* a, const double* jrk, const double* v,
                      const double* x, const double* snp, const double* crk, const double* eps);
  int start_grav_calc(int nj, int ni, const int* id, const double (*xi)[3], const double (*vi)[3],
                      const double (*ai)[3], double eps2, const double* h2);
  int get_grav_results(int nj, int ni, double (*acc)[3], double (*jrk)[3], double (*snp)[3],
                       double (*crk)[3], double* pot, int* nnb, double* dsmin, bool ngb);
  // Ensembles of independent clusters (no reference counterpart).
  int set_clusters(int n, const int* nj_per_cluster);
  int cluster_offset(int c) const;
  int num_clusters() const { return (int)cl_off_.size(); }
  int start_grav_calc_clusters(int ni, const int* cluster, const int* id, const double (*xi)[3],
                               const double (*vi)[3], const double (*ai)[3], const double* eps2_per_cluster,
                               const double* h2);
  // A force call on device-resident inputs and outputs that allocates nothing
  // and never waits on the host: the call descriptor and the partial sums live
  // in the device region [scratch, scratch + bytes) (sapamd_grav_calc_noalloc).
  // ibuf_d: [xi | vi | (ai, SIXTH)] (ni x 3 each); out_d: out_fields x ni
  // doubles, field-major; no neighbour lists.  Work is queued on s.
  int grav_calc_scratch(void* scratch, size_t bytes, int nj, int ni, const int* id_d, const double* ibuf_d,
                        double eps2, double* out_d, hipStream_t s);
  int read_ngb_list();
  int get_ngb_list(int ipipe, int maxlen, int* n, int* list);
  int force_prediction(int nj);
  int retrieve_predicted_j_particle(int addr, double* mass, double* id, double* eps2, double pos[3],
                                    double vel[3], double acc[3]);
  // Array form of retrieve_predicted_j_particle for n addresses: the rows are
  // gathered on the device and only they are copied back (no counterpart in the
  // reference, whose read-back copies every j, sapporohostclass.cpp:980-995).
  int retrieve_predicted_j_particles(int n, const int* addr, double* mass, double (*pos)[3], double (*vel)[3],
                                     double (*acc)[3]);
  // Checkpoint of the j-store (no counterpart in the reference): every stored
  // field, the predicted state, the block time, a pending prediction and the
  // cluster layout, in a caller-owned host buffer.  state_bytes flushes staged
  // updates and returns the size save_state needs; load_state needs an engine
  // opened with the same order and precision and replaces its whole store.
  long long state_bytes();
  int save_state(void* buf, long long bytes);
  int load_state(const void* buf, long long bytes);
  int retrieve_j_particle_state(int addr, double* mass, double* id, double* eps2, double pos[3],
                                double vel[3], double acc[3], double jrk[3], double ppos[3],
                                double pvel[3], double pacc[3]);

  // Launch-shape control and instrumentation (not part of the reference API).
  void set_force_override(int P, int S) { ovr_P_ = P; ovr_S_ = S; }
  void set_split_major(bool on) { ovr_smajor_ = on ? 1 : 0; }
  void set_ngb_pipeline(bool on) { ovr_pipe_ = on ? 1 : 0; }
  sapk::ForceConfig choose_config(int ni, int nj, bool ngb) const;
  // Launch-shape table written by sapporo2_amd/autotune.py: lines
  // "order precision ngb log2(ni bucket) log2(nj) P log2(j per split)".
  int load_tuning(const char* path);
  int tuning_entries() const { return (int)tune_.size(); }
  const char* tuning_path() const { return tune_path_.c_str(); }
  void set_timing(bool on) { timing_ = on; }
  // Capture scatter..finalize as a hipGraph per launch shape and replay it.
  void set_graphs(bool on) { graphs_ = on; }
  bool graphs() const { return graphs_; }
  int graph_count() const { return (int)graph_cache_.size(); }
  float last_force_ms();          // force kernel time of the last call (timing on)
  float last_predict_ms();        // predictor time of the last call (timing on, 0 if none)
  // scatter (j-update) kernel time of the last call and its update count (timing on, 0 if none)
  float last_scatter_ms(int* nupd);
  float last_finalize_ms();       // finalize (reduction, list sort, host writes) time of the last call (timing on)
  float last_device_ms();         // scatter..finalize (lists packed) device time of the last call (timing on)
  sapk::ForceConfig last_config() const { return last_cfg_; }
  int device() const { return device_; }
  int num_cus() const { return cus_; }
  hipStream_t stream() const { return stream_; }
  const char* last_error() const { return err_.c_str(); }
  int synchronize();

 private:
  int fail(const char* what, hipError_t e = hipSuccess);
  int ensure_store(int need);
  int ensure_partials(size_t entries);
  int flush_updates(int ni_reset, int pred_j1 = 0);
  int upload_updates(int* nupd);
  int copy_staged();
  int scatter_updates(int n, int ni_reset, int pred_j1 = 0);
  int predict_range(int j0, int j1);
  int post_desc(int ni, int nupd, int ni_reset, double eps2);
  bool pack_i(int i0, int i1, const int* id, const double (*xi)[3], const double (*vi)[3], const double (*ai)[3],
              const double* h2, double* bx, double* bv, double* ba, double* bh, int* bid);
  int scatter_blocks(int nupd, int ni_reset) const;
  int enqueue_call(int grid_ni, int nj, const sapk::ForceConfig& cfg, bool ngb_on, bool predict,
                   int sblocks, bool rows);
  int launch_call_graph(int grid_ni, int nj, const sapk::ForceConfig& cfg, bool ngb_on, bool predict,
                        int sblocks, bool rows);
  void drop_graphs();
  int wait_results();
  int ensure_lists(int need);
  int ensure_packed(long need);
  sapk::FinalizeArgs finalize_args(int grid_ni, bool ngb_on) const;
  const double* result_field(int f) const { return res_h_ + (size_t)f * res_stride_; }
  int fetch_j_host(int nj);
  int grow_staging(int need);
  int ensure_icap(int need);

  bool open_ = false;
  int device_ = 0, order_ = sapk::ORDER_H4, cus_ = 256;
  bool f64_ = false;
  hipStream_t stream_ = nullptr;
  std::string err_;

  // j-store
  sapk::StoreView st_{};
  int nj_max_ = NJ_INITIAL;   // logical capacity (reference growth rule)
  int max_addr_ = -1;         // largest address ever stored

  // staged updates (pinned), double-buffered
  struct Stage {
    double* d = nullptr;      // cap rows of nf doubles + one (address, id) int pair
    int cap = 0;
    hipEvent_t ev = nullptr;  // completion of the copy that last read it
    bool busy = false;
    uint64_t seq = 0;         // position of ev among the stream's recorded events
  };
  Stage stage_[2];
  int cur_ = 0, nstaged_ = 0;
  int flushed_ = 0;           // staged rows [0, flushed_) already copied to the device (early flush)
  std::vector<int> slot_of_;  // address -> staged slot or -1
  std::vector<int> staged_addr_;  // addresses of the staged rows, in row order
  double* upd_d_ = nullptr;
  int upd_cap_d_ = 0;

  // prediction state
  double t_i_ = 0.0;
  bool predict_ = false;

  // i-particles and results
  double* ibuf_h_ = nullptr; int* iid_h_ = nullptr;
  double* ibuf_d_ = nullptr; int* iid_d_ = nullptr;
  double* part_d_ = nullptr; double* pds2_d_ = nullptr; int* paddr_d_ = nullptr;
  size_t part_cap_ = 0;
  // host-mapped pinned results and packed lists (host pointer, device alias)
  double* res_h_ = nullptr; double* res_hd_ = nullptr; long res_stride_ = 0;
  int* loff_h_ = nullptr; int* loff_hd_ = nullptr;
  int* packed_h_ = nullptr; int* packed_hd_ = nullptr; long packed_cap_ = 0;
  int* pack_ctr_d_ = nullptr;
  int* cnt_d_ = nullptr; int* list_d_ = nullptr; int list_cap_ = 0;
  std::vector<int> slot_lists_h_;   // device list slots read back when the packed lists did not fit
  bool lists_in_slots_ = false;     // lists read back from the device slots (packed overflow)
  bool lists_packed_ = false;       // last call's host lists are atomically packed (ensembles), else in i order
  hipEvent_t ev_done_ = nullptr, ev_d0_ = nullptr, ev_d1_ = nullptr, ev_f0_ = nullptr, ev_f1_ = nullptr, ev_p0_ = nullptr, ev_p1_ = nullptr;
  hipEvent_t ev_s0_ = nullptr, ev_s1_ = nullptr;
  bool timed_scatter_ = false;
  int last_nupd_ = 0;
  bool inflight_ = false, have_results_ = false, ngb_fetched_ = false, ngb_on_last_ = false;
  bool timed_predict_ = false, timed_force_ = false;
  bool capturing_ = false;
  int ni_last_ = 0, nj_last_ = 0;
  sapk::ForceConfig last_cfg_{64, 1, 1, 0, 0};
  bool timing_ = false;
  int ovr_P_ = 0, ovr_S_ = 0, ovr_smajor_ = 0, ovr_pipe_ = 0;

  // per-call descriptor (pinned host copy + device copy) and graph cache
  double* call_h_ = nullptr;           // pinned call block: descriptor | i ids | i doubles
  double* call_d_ = nullptr;           // its device copy (one host->device copy per call)
  const double* call_ibuf_d_ = nullptr;  // i doubles and ids of the current call in call_d_
  const int* call_iid_d_ = nullptr;
  sapk::CallDesc* desc_h_ = nullptr;   // the descriptor slot at the start of the call block
  sapk::CallDesc* desc_d_ = nullptr;
  hipEvent_t ev_desc_ = nullptr;
  bool desc_pending_ = false;
  uint64_t desc_seq_ = 0;
  // recorded events are numbered in stream order; done_seq_ is the number of the
  // latest completion event seen complete (everything recorded before it is done)
  uint64_t rec_seq_ = 0, done_seq_ = 0, done_rec_seq_ = 0;
  bool graphs_ = false;
  uint64_t gen_ = 1;          // bumped whenever a buffer a captured graph uses is reallocated
  uint64_t tick_ = 0;
  struct GraphKey {
    int grid_ni, nj, P, S, jchunk, smajor, pipe, ngb_on, predict, sblocks;
    uint64_t gen;
    bool operator==(const GraphKey& o) const {
      return grid_ni == o.grid_ni && nj == o.nj && P == o.P && S == o.S && jchunk == o.jchunk &&
             smajor == o.smajor && pipe == o.pipe && ngb_on == o.ngb_on && predict == o.predict &&
             sblocks == o.sblocks && gen == o.gen;
    }
  };
  struct GraphEntry { GraphKey key; hipGraphExec_t exec; uint64_t used; };
  std::vector<GraphEntry> graph_cache_;

  // ensembles
  int icap_ = 0;
  std::vector<int> cl_off_, cl_n_;
  int cl_total_ = 0;
  // work table of an ensemble call: [items | launch order | groups], built in
  // pinned host memory and uploaded with one copy
  char* ens_tab_h_ = nullptr; char* ens_tab_d_ = nullptr; size_t ens_tab_cap_ = 0;
  int4* idesc_d_ = nullptr;
  double* ens_eps_d_ = nullptr; int ens_eps_cap_ = 0;

  struct TuneEntry { int order, prec, ngb, li, lj, P, jlog; };
  std::vector<TuneEntry> tune_;
  std::string tune_path_, arch_;

  // read-back cache of the j-store (predicted and raw state)
  bool jhost_valid_ = false;
  int jhost_n_ = 0;
  std::vector<double> jpred_h_, jraw_h_;
  std::vector<int> jid_h_;
  double* gather_d_ = nullptr; size_t gather_cap_ = 0;
  int* gaddr_d_ = nullptr; size_t gaddr_cap_ = 0;
  std::vector<double> gather_h_;
};

}  // namespace sap

// The object behind a sapamd_engine handle (include/sapporo_amd.h); the g6
// and 6th-order ABIs hold one as their global instance.
struct sapamd_engine {
  sap::Engine eng;
};
