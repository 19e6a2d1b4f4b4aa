// engine.cpp -- host engine (see engine.h).  Reference behaviour cited per method.
#include "engine.h"
#include "hostpool.h"

#include <dlfcn.h>
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

namespace sap {

using namespace sapk;

#define HIPCHK(call, what)                       \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return fail(what, e_); \
  } while (0)

namespace {
constexpr int WAVES_PER_SIMD_TARGET = 32;   // several dispatch rounds balance better
constexpr int MIN_ITERS_PER_LANE = 16;
constexpr int MIN_J_PER_SPLIT = 512;
constexpr int FOLD = 64;   // interactions per lane between FP64 folds (kernels.hip KT)
constexpr int IBUF_FIELDS = 10;  // xi 3, vi 3, ai 3, h2 1
// Staged j-update rows are copied to the device as soon as this many are
// pending, while the caller is still staging (the copy overlaps its loop).
constexpr int EARLY_FLUSH_ROWS = 4096;
// Upload block of a single-system call, one host->device copy: the call
// descriptor, then the i ids, then the i doubles.  The ids region is sized by
// the power-of-two bucket of ni, so every offset depends on the bucket only
// (the pointers a captured graph holds stay valid for its bucket).
constexpr size_t CALL_HDR = 256;
inline size_t call_ids_bytes(int bucket) { return ((size_t)bucket * sizeof(int) + 255) / 256 * 256; }
constexpr size_t CALL_DOUBLES = (CALL_HDR + ((size_t)NPIPES * sizeof(int) + 255) / 256 * 256) / sizeof(double) +
                                (size_t)IBUF_FIELDS * NPIPES;

template <typename T>
hipError_t dmalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}
template <typename T>
hipError_t hmalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipHostMalloc(reinterpret_cast<void**>(p), n * sizeof(T), hipHostMallocDefault);
}
// Pinned host memory mapped into the device's address space: the kernels
// write results into it directly (host pointer h, device alias d).  Coherent
// (fine-grained) memory is written through to the host during the kernel;
// non-coherent memory is cached in the GPU's L2 and written back by the
// system-scope release of the call's completion event.  Choice:
// $SAPPORO2_AMD_HOST_COHERENT (default shown in host_coherent()).
bool host_coherent() {
  const char* v = getenv("SAPPORO2_AMD_HOST_COHERENT");
  return v ? atoi(v) != 0 : true;
}
template <typename T>
hipError_t hmap(T** h, T** d, size_t n) {
  *h = *d = nullptr;
  if (n == 0) n = 1;
  const unsigned flags = hipHostMallocMapped | (host_coherent() ? hipHostMallocCoherent : hipHostMallocNonCoherent);
  const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(h), n * sizeof(T), flags);
  if (e != hipSuccess) return e;
  return hipHostGetDevicePointer(reinterpret_cast<void**>(d), *h, 0);
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
void hfree(T*& p) {
  if (p) (void)hipHostFree(p);
  p = nullptr;
}
}  // namespace

Engine::Engine() = default;
Engine::~Engine() {
  if (open_) close();
}

int Engine::fail(const char* what, hipError_t e) {
  err_ = what;
  if (e != hipSuccess) {
    err_ += ": ";
    err_ += hipGetErrorString(e);
  }
  fprintf(stderr, "sapporo2_amd: %s\n", err_.c_str());
  return -1;
}

// Reference sapporo::open (sapporohostclass.cpp:90-170): one device, kernels
// are compiled in, initial j capacity 16384.
int Engine::open(int device, int order, int precision) {
  if (open_) close();
  if (order != ORDER_G5 && order != ORDER_H4 && order != ORDER_H6) return fail("unsupported integration order");
  if (precision != PREC_DS && precision != PREC_DOUBLE) return fail("unsupported precision");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (ndev <= 0) return fail("no HIP device");
  if (device < 0 || device >= ndev) device = 0;
  device_ = device;
  order_ = order;
  f64_ = precision == PREC_DOUBLE;
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
  cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  arch_ = prop.gcnArchName;
  arch_ = arch_.substr(0, arch_.find(':'));
  {
    // tuned launch shapes: $SAPPORO2_AMD_TUNE, else tune/<arch>.txt beside the library
    std::string path;
    if (const char* env = getenv("SAPPORO2_AMD_TUNE")) {
      path = env;
    } else {
      Dl_info info;
      if (dladdr(reinterpret_cast<void*>(&sapk::launch_force), &info) && info.dli_fname) {
        path = info.dli_fname;
        path = path.substr(0, path.rfind('/') + 1) + "tune/" + arch_ + ".txt";
      }
    }
    if (!path.empty()) load_tuning(path.c_str());
  }
  HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  HIPCHK(hipEventCreateWithFlags(&ev_done_, hipEventDisableTiming), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_d0_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_d1_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_f0_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_f1_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_p0_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_p1_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_s0_), "hipEventCreate");
  HIPCHK(hipEventCreate(&ev_s1_), "hipEventCreate");
  for (auto& s : stage_) HIPCHK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "hipEventCreate");
  HIPCHK(hipEventCreateWithFlags(&ev_desc_, hipEventDisableTiming), "hipEventCreate");
  HIPCHK(dmalloc(&pack_ctr_d_, 1), "pack counter");
  HIPCHK(hmalloc(&call_h_, CALL_DOUBLES), "pinned call block");
  HIPCHK(dmalloc(&call_d_, CALL_DOUBLES), "call block");
  desc_h_ = reinterpret_cast<CallDesc*>(call_h_);
  desc_d_ = reinterpret_cast<CallDesc*>(call_d_);
  desc_pending_ = false;
  rec_seq_ = done_seq_ = done_rec_seq_ = desc_seq_ = 0;
  flushed_ = 0;
  if (const char* g = getenv("SAPPORO2_AMD_GRAPHS")) graphs_ = atoi(g) != 0;

  icap_ = 0;
  if (ensure_icap(NPIPES) != 0) return -1;

  nj_max_ = NJ_INITIAL;
  max_addr_ = -1;
  slot_of_.assign(nj_max_, -1);
  staged_addr_.clear();
  nstaged_ = 0;
  cur_ = 0;
  predict_ = false;
  t_i_ = 0.0;
  open_ = true;
  if (ensure_store(NJ_INITIAL) != 0) return -1;
  if (grow_staging(1024) != 0) return -1;
  HIPCHK(hipStreamSynchronize(stream_), "open sync");
  return 0;
}

// Reference sapporo::close (sapporohostclass.cpp:391-402).
int Engine::close() {
  if (!open_) return 0;
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  drop_graphs();
  hfree(call_h_); dfree(call_d_);
  desc_h_ = nullptr; desc_d_ = nullptr;
  if (ev_desc_) (void)hipEventDestroy(ev_desc_);
  ev_desc_ = nullptr;
  desc_pending_ = false;
  dfree(st_.raw); dfree(st_.id); dfree(st_.rec); dfree(st_.ppos); dfree(st_.pvel); dfree(st_.pacc); dfree(st_.idr);
  st_.cap = 0;
  for (auto& s : stage_) {
    hfree(s.d); s.cap = 0; s.busy = false;
    if (s.ev) (void)hipEventDestroy(s.ev);
    s.ev = nullptr;
  }
  dfree(upd_d_); upd_cap_d_ = 0;
  hfree(ibuf_h_); hfree(iid_h_); dfree(ibuf_d_); dfree(iid_d_);
  dfree(part_d_); dfree(pds2_d_); dfree(paddr_d_); part_cap_ = 0;
  hfree(res_h_); res_hd_ = nullptr; res_stride_ = 0;
  hfree(loff_h_); loff_hd_ = nullptr;
  hfree(packed_h_); packed_hd_ = nullptr; packed_cap_ = 0;
  dfree(pack_ctr_d_);
  dfree(cnt_d_); dfree(list_d_); list_cap_ = 0;
  slot_lists_h_.clear();
  dfree(gather_d_); gather_cap_ = 0;
  dfree(gaddr_d_); gaddr_cap_ = 0;
  dfree(idesc_d_); dfree(ens_eps_d_); ens_eps_cap_ = 0;
  hfree(ens_tab_h_); dfree(ens_tab_d_); ens_tab_cap_ = 0;
  icap_ = 0; cl_off_.clear(); cl_n_.clear(); cl_total_ = 0;
  for (hipEvent_t* e : {&ev_done_, &ev_d0_, &ev_d1_, &ev_f0_, &ev_f1_, &ev_p0_, &ev_p1_, &ev_s0_, &ev_s1_}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  open_ = inflight_ = have_results_ = ngb_fetched_ = jhost_valid_ = false;
  slot_of_.clear();
  staged_addr_.clear();
  nstaged_ = 0;
  return 0;
}

int Engine::synchronize() {
  if (!open_) return 0;
  HIPCHK(hipStreamSynchronize(stream_), "synchronize");
  inflight_ = false;
  return 0;
}

// Device j-store growth.  New slots are zeroed (massless, at the origin).
int Engine::ensure_store(int need) {
  if (need <= st_.cap) return 0;
  const int oldcap = st_.cap;
  long nc = std::max<long>({(long)need, (long)oldcap + oldcap / 2, (long)NJ_INITIAL});
  nc = (nc + 255) / 256 * 256;  // a multiple of JTILE and of the predictor's block
  const int newcap = (int)nc;
  const int nfr = raw_fields(order_);
  const size_t rb = jrec_bytes(order_, f64_);
  StoreView ns{};
  ns.cap = newcap;
  HIPCHK(dmalloc(&ns.raw, (size_t)nfr * newcap), "j-store raw");
  HIPCHK(dmalloc(&ns.id, (size_t)newcap), "j-store ids");
  HIPCHK(hipMalloc(&ns.rec, rb * newcap), "j-store records");
  HIPCHK(dmalloc(&ns.idr, (size_t)newcap / JTILE), "j-store tile id ranges");
  // zeroed slots carry id 0, so every new tile's range starts at [0, 0]
  HIPCHK(hipMemsetAsync(ns.idr, 0, sizeof(int2) * (size_t)newcap / JTILE, stream_), "memset");
  if (!f64_) {
    HIPCHK(dmalloc(&ns.ppos, (size_t)3 * newcap), "j-store ppos");
    HIPCHK(dmalloc(&ns.pvel, (size_t)3 * newcap), "j-store pvel");
    if (order_ >= ORDER_H6) HIPCHK(dmalloc(&ns.pacc, (size_t)3 * newcap), "j-store pacc");
  }
  HIPCHK(hipMemsetAsync(ns.raw, 0, sizeof(double) * nfr * (size_t)newcap, stream_), "memset");
  HIPCHK(hipMemsetAsync(ns.id, 0, sizeof(int) * (size_t)newcap, stream_), "memset");
  HIPCHK(hipMemsetAsync(ns.rec, 0, rb * newcap, stream_), "memset");
  if (ns.ppos) HIPCHK(hipMemsetAsync(ns.ppos, 0, sizeof(double) * 3 * (size_t)newcap, stream_), "memset");
  if (ns.pvel) HIPCHK(hipMemsetAsync(ns.pvel, 0, sizeof(double) * 3 * (size_t)newcap, stream_), "memset");
  if (ns.pacc) HIPCHK(hipMemsetAsync(ns.pacc, 0, sizeof(double) * 3 * (size_t)newcap, stream_), "memset");
  if (oldcap > 0) {
    for (int f = 0; f < nfr; ++f)
      HIPCHK(hipMemcpyAsync(ns.raw + (size_t)f * newcap, st_.raw + (size_t)f * oldcap,
                            sizeof(double) * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
    HIPCHK(hipMemcpyAsync(ns.id, st_.id, sizeof(int) * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
    HIPCHK(hipMemcpyAsync(ns.rec, st_.rec, rb * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
    HIPCHK(hipMemcpyAsync(ns.idr, st_.idr, sizeof(int2) * (size_t)oldcap / JTILE, hipMemcpyDeviceToDevice,
                          stream_), "store copy");
    for (int d = 0; d < 3; ++d) {
      if (ns.ppos)
        HIPCHK(hipMemcpyAsync(ns.ppos + (size_t)d * newcap, st_.ppos + (size_t)d * oldcap,
                              sizeof(double) * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
      if (ns.pvel)
        HIPCHK(hipMemcpyAsync(ns.pvel + (size_t)d * newcap, st_.pvel + (size_t)d * oldcap,
                              sizeof(double) * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
      if (ns.pacc)
        HIPCHK(hipMemcpyAsync(ns.pacc + (size_t)d * newcap, st_.pacc + (size_t)d * oldcap,
                              sizeof(double) * oldcap, hipMemcpyDeviceToDevice, stream_), "store copy");
    }
    HIPCHK(hipStreamSynchronize(stream_), "store grow sync");
  }
  ++gen_;
  dfree(st_.raw); dfree(st_.id); dfree(st_.rec); dfree(st_.ppos); dfree(st_.pvel); dfree(st_.pacc); dfree(st_.idr);
  st_ = ns;
  return 0;
}

int Engine::grow_staging(int need) {
  const int nr = raw_fields(order_) + 1;  // row: fields, then the (address, id) pair
  for (auto& s : stage_) {
    if (s.cap >= need) continue;
    if (s.busy && s.seq > done_seq_) HIPCHK(hipEventSynchronize(s.ev), "staging wait");
    s.busy = false;
    const int nc = std::max(need, s.cap * 2);
    double* d = nullptr;
    HIPCHK(hmalloc(&d, (size_t)nc * nr), "pinned staging");
    if (s.cap > 0 && &s == &stage_[cur_] && nstaged_ > 0) memcpy(d, s.d, sizeof(double) * (size_t)nstaged_ * nr);
    hfree(s.d);
    s.d = d;
    s.cap = nc;
  }
  return 0;
}

// Reference set_time / set_no_time (sapporohostclass.cpp:405-427).
int Engine::set_time(double t) {
  t_i_ = t;
  predict_ = order_ != ORDER_G5;  // second order has no predictor (kernels.cu:326)
  jhost_valid_ = false;
  return 0;
}
int Engine::set_no_time() {
  predict_ = false;
  jhost_valid_ = false;
  return 0;
}

// Reference set_j_particle (sapporohostclass.cpp:434-558): stage the update;
// re-setting an address before the next flush overwrites its slot; the
// logical capacity grows by 10% and addresses beyond that are rejected.
int Engine::set_j_particle(int addr, int id, double tj, double dtj, double mass, const double a[3],
                           const double jrk[3], const double v[3], const double x[3],
                           const double snp[3], const double crk[3], double eps) {
  if (!open_) return fail("set_j_particle: library not open");
  if (addr < 0) return fail("set_j_particle: negative address");
  if (addr >= nj_max_) {
    fprintf(stderr, "sapporo2_amd: increasing nj_max from %d for address %d\n", nj_max_, addr);
    nj_max_ = (int)(nj_max_ * 1.1) + 1;
    if (addr >= nj_max_) {
      fprintf(stderr, "sapporo2_amd: address %d beyond the grown capacity %d; send particles in order\n",
              addr, nj_max_);
      return fail("set_j_particle: address out of order");
    }
  }
  if ((int)slot_of_.size() < nj_max_) slot_of_.resize(nj_max_, -1);
  Stage* sg = &stage_[cur_];
  // a stage still read by the copy of an earlier flush is waited for before its
  // first row is rewritten; copies of this flush's own rows (early flushes)
  // need no wait: a rewritten row is copied again behind them
  if (nstaged_ == 0 && sg->busy) {
    if (sg->seq > done_seq_) HIPCHK(hipEventSynchronize(sg->ev), "staging wait");
    sg->busy = false;
  }
  int k = slot_of_[addr];
  if (k < 0) {
    if (nstaged_ >= sg->cap) {
      if (grow_staging(nstaged_ + 1) != 0) return -1;
      sg = &stage_[cur_];
    }
    k = nstaged_++;
    slot_of_[addr] = k;
    staged_addr_.push_back(addr);
  } else if (k < flushed_) {
    flushed_ = k;  // already copied: copy again from this row
  }
  const int nf = raw_fields(order_);
  double* u = sg->d + (size_t)k * (nf + 1);
  u[RF_T] = tj; u[RF_DT] = dtj; u[RF_M] = mass; u[RF_EPS] = eps;
  for (int d = 0; d < 3; ++d) {
    u[RF_X + d] = x ? x[d] : 0.0;
    u[RF_VX + d] = v ? v[d] : 0.0;
    u[RF_AX + d] = a ? a[d] : 0.0;
    u[RF_JX + d] = jrk ? jrk[d] : 0.0;
    if (order_ >= ORDER_H6) {
      u[RF_SX + d] = snp ? snp[d] : 0.0;
      u[RF_CX + d] = crk ? crk[d] : 0.0;
    }
  }
  const int pair[2] = {addr, id};
  memcpy(u + nf, pair, sizeof(pair));
  if (addr > max_addr_) max_addr_ = addr;
  jhost_valid_ = false;
  if (nstaged_ - flushed_ >= EARLY_FLUSH_ROWS) return copy_staged();
  return 0;
}

// Copies the staged rows [flushed_, nstaged_) of the current stage to the
// device staging buffer (asynchronously, behind earlier work on the stream).
// The device buffer only grows; after a regrowth every staged row is copied.
int Engine::copy_staged() {
  const int n = nstaged_;
  if (n <= flushed_) return 0;
  const int nr = raw_fields(order_) + 1;
  if (upd_cap_d_ < n) {
    if (stream_) HIPCHK(hipStreamSynchronize(stream_), "staging grow sync");
    dfree(upd_d_);
    const int nc = std::max(n, upd_cap_d_ * 2);
    HIPCHK(dmalloc(&upd_d_, (size_t)nc * nr), "device staging");
    upd_cap_d_ = nc;
    flushed_ = 0;
    ++gen_;
  }
  Stage& sg = stage_[cur_];
  const size_t o = (size_t)flushed_ * nr;
  HIPCHK(hipMemcpyAsync(upd_d_ + o, sg.d + o, sizeof(double) * ((size_t)n * nr - o), hipMemcpyHostToDevice,
                        stream_), "update upload");
  HIPCHK(hipEventRecord(sg.ev, stream_), "event");
  sg.busy = true;
  sg.seq = ++rec_seq_;
  flushed_ = n;
  return 0;
}

int Engine::set_j_particles(int n, const int* addr, const int* id, const double* tj, const double* dtj,
                            const double* mass, const double* a, const double* jrk, const double* v,
                            const double* x, const double* snp, const double* crk, const double* eps) {
  if (n < 0 || !addr || !id || !mass || !x) return fail("set_j_particles: bad arguments");
  if (!open_) return fail("set_j_particles: library not open");
  // the batch names all its addresses up front: grow the logical capacity once
  int amax = -1;
  for (int k = 0; k < n; ++k) amax = addr[k] > amax ? addr[k] : amax;
  if (amax >= nj_max_) nj_max_ = amax + 1;
  for (int k = 0; k < n; ++k) {
    const size_t o = 3 * (size_t)k;
    const int rc = set_j_particle(addr[k], id[k], tj ? tj[k] : 0.0, dtj ? dtj[k] : 0.0, mass[k],
                                  a ? a + o : nullptr, jrk ? jrk + o : nullptr, v ? v + o : nullptr, x + o,
                                  snp ? snp + o : nullptr, crk ? crk + o : nullptr, eps ? eps[k] : 0.0);
    if (rc != 0) return rc;
  }
  return 0;
}

// Reference send_j_particles_to_device + copyJInDev (sapporohostclass.cpp:889-923,
// 1204-1248), fused with the neighbour-counter reset of dev_reset_buffers.
// upload_updates issues the host->device copy of the staged updates; the
// scatter kernel is launched separately so callers can bracket device work.
int Engine::upload_updates(int* nupd) {
  const int n = nstaged_;
  *nupd = n;
  if (n <= 0) return 0;
  if (ensure_store(max_addr_ + 1) != 0) return -1;
  if (copy_staged() != 0) return -1;
  for (const int a : staged_addr_) slot_of_[a] = -1;
  staged_addr_.clear();
  nstaged_ = 0;
  flushed_ = 0;
  cur_ ^= 1;
  return 0;
}

// Work-groups of the scatter launch: grid-stride over the descriptor's counts,
// rounded up to a power of two so that calls of similar size share a shape.
int Engine::scatter_blocks(int n, int ni_reset) const {
  const long m = std::max(n, ni_reset);
  if (m <= 0) return 0;
  int nb = 1;
  while ((long)nb * 256 < m && nb < 4096) nb <<= 1;
  return nb;
}

int Engine::scatter_updates(int n, int ni_reset, int pred_j1) {
  ScatterArgs a{};
  a.st = st_;
  a.upd = upd_d_;
  a.nf = raw_fields(order_);
  a.ngb_cnt = cnt_d_;
  a.pack_ctr = pack_ctr_d_;
  a.cd = desc_d_;
  a.grid_blocks = scatter_blocks(n, ni_reset);
  a.pred_j1 = pred_j1;
  if (a.grid_blocks == 0) return 0;
  if (launch_scatter(stream_, order_, f64_, a) != 0) return fail("scatter launch", hipGetLastError());
  return 0;
}

// Writes the call descriptor (pinned host copy) and queues its upload.  The
// previous upload has to have read the host copy before it is rewritten.
int Engine::post_desc(int ni, int nupd, int ni_reset, double eps2) {
  if (desc_pending_ && desc_seq_ > done_seq_) HIPCHK(hipEventSynchronize(ev_desc_), "descriptor wait");
  desc_h_->ti = t_i_;
  desc_h_->eps2 = eps2;
  desc_h_->ni = ni;
  desc_h_->nupd = nupd;
  desc_h_->ni_reset = ni_reset;
  desc_h_->pad = 0;
  HIPCHK(hipMemcpyAsync(desc_d_, desc_h_, sizeof(CallDesc), hipMemcpyHostToDevice, stream_), "descriptor upload");
  HIPCHK(hipEventRecord(ev_desc_, stream_), "event");
  desc_pending_ = true;
  desc_seq_ = ++rec_seq_;
  return 0;
}

int Engine::flush_updates(int ni_reset, int pred_j1) {
  int n = 0;
  if (upload_updates(&n) != 0) return -1;
  if (post_desc(0, n, ni_reset, 0.0) != 0) return -1;
  return scatter_updates(n, ni_reset, pred_j1);
}

int Engine::predict_range(int j0, int j1) {
  if (j1 <= j0) return 0;
  if (ensure_store(j1) != 0) return -1;
  PredictArgs p{};
  p.st = st_;
  p.j0 = j0;
  p.j1 = j1;
  p.cd = desc_d_;
  const bool ev = timing_ && !capturing_;
  if (ev) HIPCHK(hipEventRecord(ev_p0_, stream_), "event");
  if (launch_predict(stream_, order_, f64_, p) != 0) return fail("predict launch", hipGetLastError());
  if (ev) HIPCHK(hipEventRecord(ev_p1_, stream_), "event");
  timed_predict_ = ev;
  return 0;
}

int Engine::ensure_partials(size_t entries) {
  if (entries <= part_cap_) return 0;
  dfree(part_d_);
  dfree(pds2_d_);
  dfree(paddr_d_);
  const size_t nc = std::max(entries, part_cap_ + part_cap_ / 2);
  HIPCHK(dmalloc(&part_d_, nc * part_fields(order_)), "partials");
  HIPCHK(dmalloc(&pds2_d_, nc), "partials");
  HIPCHK(dmalloc(&paddr_d_, nc), "partials");
  part_cap_ = nc;
  ++gen_;
  return 0;
}

// Launch shape: P i-particles per wave (64 unless fewer are active), S j-splits
// so that about WAVES_PER_SIMD_TARGET waves per SIMD exist, each split at least
// MIN_ITERS_PER_LANE iterations per lane.
int Engine::load_tuning(const char* path) {
  tune_.clear();
  tune_path_.clear();
  FILE* f = path ? fopen(path, "r") : nullptr;
  if (!f) return 0;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    TuneEntry t{};
    if (sscanf(line, "%d %d %d %d %d %d %d", &t.order, &t.prec, &t.ngb, &t.li, &t.lj, &t.P, &t.jlog) == 7 &&
        force_p_supported(t.P) && t.jlog >= 0 && t.jlog < 31)
      tune_.push_back(t);
  }
  fclose(f);
  tune_path_ = path;
  return (int)tune_.size();
}

static int ceil_log2(long v) {
  int l = 0;
  while ((1L << l) < v) ++l;
  return l;
}

// Launch shape: the tuned table's entry for (order, precision, neighbour lists,
// active-count bucket, nearest nj) when present; otherwise P i-particles per
// wave (64 unless fewer are active) and S j-splits for ~WAVES_PER_SIMD_TARGET
// waves per SIMD with at least MIN_J_PER_SPLIT j per split.
ForceConfig Engine::choose_config(int ni, int nj, bool ngb) const {
  int P = 64;
  if (ni < 64) {
    P = 1;
    while (P < ni) P <<= 1;
  }
  long jchunk_t = 0;
  if (!tune_.empty() && ovr_P_ <= 0 && ovr_S_ <= 0) {
    const int li = ceil_log2(ni), lj = ceil_log2(nj);
    const int prec = f64_ ? PREC_DOUBLE : PREC_DS;
    const TuneEntry* best = nullptr;
    for (const TuneEntry& t : tune_)
      if (t.order == order_ && t.prec == prec && t.ngb == (ngb ? 1 : 0) && t.li == li &&
          (!best || std::abs(t.lj - lj) < std::abs(best->lj - lj)))
        best = &t;
    if (best) {
      P = best->P;
      jchunk_t = 1L << best->jlog;
    }
  }
  if (ovr_P_ > 0 && force_p_supported(ovr_P_)) P = ovr_P_;
  const int Q = 64 / P;
  const long G = (ni + P - 1) / P;
  long S;
  if (jchunk_t > 0) {
    S = (nj + jchunk_t - 1) / jchunk_t;
  } else {
    const long target = (long)cus_ * 4 * WAVES_PER_SIMD_TARGET;
    S = (target + G - 1) / G;
    const long minj = std::max<long>((long)Q * MIN_ITERS_PER_LANE, MIN_J_PER_SPLIT);
    const long maxS = std::max<long>(1, (nj + minj - 1) / minj);
    S = std::min(S, maxS);
  }
  if (ovr_S_ > 0) S = ovr_S_;
  if (S < 1) S = 1;
  long jchunk = (nj + S - 1) / S;
  const long tile = (long)Q * FOLD;
  if (jchunk >= tile) jchunk = (jchunk + tile - 1) / tile * tile;
  else jchunk = (jchunk + Q - 1) / Q * Q;
  if (jchunk < 1) jchunk = 1;
  S = (nj + jchunk - 1) / jchunk;
  if (S < 1) S = 1;
  return ForceConfig{P, (int)S, (int)jchunk, ovr_smajor_, ovr_pipe_};
}

// i-particle-sized buffers (ensemble inputs, host-mapped result rows and list
// offsets, neighbour counts) for up to `need` i-particles per call.  The
// neighbour-list slots and the packed host lists are sized separately, only
// for calls with lists (ensure_lists, ensure_packed).
int Engine::ensure_icap(int need) {
  if (need <= icap_) return 0;
  if (stream_) HIPCHK(hipStreamSynchronize(stream_), "i-buffer grow sync");
  inflight_ = false;
  hfree(ibuf_h_); hfree(iid_h_); dfree(ibuf_d_); dfree(iid_d_);
  hfree(res_h_); hfree(loff_h_); dfree(cnt_d_); dfree(idesc_d_);
  const size_t nI = (size_t)need;
  HIPCHK(hmalloc(&ibuf_h_, IBUF_FIELDS * nI), "pinned i buffer");
  HIPCHK(hmalloc(&iid_h_, nI), "pinned i ids");
  HIPCHK(dmalloc(&ibuf_d_, IBUF_FIELDS * nI), "i buffer");
  HIPCHK(dmalloc(&iid_d_, nI), "i ids");
  HIPCHK(hmap(&res_h_, &res_hd_, nI * out_fields(order_)), "host-mapped results");
  res_stride_ = (long)nI;
  HIPCHK(hmap(&loff_h_, &loff_hd_, nI), "host-mapped list offsets");
  HIPCHK(dmalloc(&cnt_d_, nI), "neighbour counts");
  HIPCHK(dmalloc(&idesc_d_, nI), "ensemble i descriptors");
  HIPCHK(hipMemsetAsync(cnt_d_, 0, nI * sizeof(int), stream_), "memset");
  icap_ = need;
  have_results_ = false;
  ++gen_;
  return 0;
}

// Device neighbour-list slots (NGB_CAP per i-particle) for `need` i-particles.
int Engine::ensure_lists(int need) {
  if (need <= list_cap_) return 0;
  if (stream_) HIPCHK(hipStreamSynchronize(stream_), "list grow sync");
  inflight_ = false;
  dfree(list_d_);
  HIPCHK(dmalloc(&list_d_, (size_t)need * NGB_CAP), "neighbour lists");
  list_cap_ = need;
  have_results_ = false;
  ++gen_;
  return 0;
}

// Host-mapped packed lists of at least `need` ids.  A call whose lists do not
// fit is still correct: read_ngb_list then copies the device slots back.
int Engine::ensure_packed(long need) {
  if (need <= packed_cap_) return 0;
  if (stream_) HIPCHK(hipStreamSynchronize(stream_), "packed-list grow sync");
  inflight_ = false;
  hfree(packed_h_);
  HIPCHK(hmap(&packed_h_, &packed_hd_, (size_t)need), "host-mapped packed lists");
  packed_cap_ = need;
  have_results_ = false;
  ++gen_;
  return 0;
}

FinalizeArgs Engine::finalize_args(int grid_ni, bool ngb_on) const {
  FinalizeArgs fz{};
  fz.part = part_d_;
  fz.part_ds2 = pds2_d_;
  fz.part_addr = paddr_d_;
  fz.cd = desc_d_;
  fz.grid_ni = grid_ni;
  fz.rec = st_.rec;
  fz.ngb_cnt = cnt_d_;
  fz.ngb_list = list_d_;
  fz.out = res_hd_;
  fz.ostride = res_stride_;
  fz.list_off = loff_hd_;
  fz.packed = packed_hd_;
  fz.packed_cap = packed_cap_;
  fz.pack_ctr = pack_ctr_d_;
  fz.inorder = 1;
  fz.ngb_on = ngb_on ? 1 : 0;
  return fz;
}

// Ensemble layout: cluster c owns the addresses [offset(c), offset(c) + nj_c),
// offsets aligned to the JTILE tile.  The reference has no counterpart (its
// GRAPE-6 ABI ignores cluster_id, sapporoG6lib.cpp:99).
int Engine::set_clusters(int n, const int* njc) {
  if (!open_) return fail("set_clusters: library not open");
  if (n <= 0 || !njc) return fail("set_clusters: bad arguments");
  cl_off_.assign(n, 0);
  cl_n_.assign(n, 0);
  long off = 0;
  for (int c = 0; c < n; ++c) {
    if (njc[c] < 0) return fail("set_clusters: negative cluster size");
    cl_off_[c] = (int)off;
    cl_n_[c] = njc[c];
    off += (njc[c] + JTILE - 1) / JTILE * JTILE;
    if (off > (1L << 30)) return fail("set_clusters: too many particles");
  }
  nj_max_ = std::max(nj_max_, (int)off);
  if ((int)slot_of_.size() < nj_max_) slot_of_.resize(nj_max_, -1);
  if (ensure_store((int)off) != 0) return -1;
  cl_total_ = (int)off;
  return 0;
}

int Engine::cluster_offset(int c) const {
  return (c >= 0 && c < (int)cl_off_.size()) ? cl_off_[c] : -1;
}

// Forces for i-particles of many clusters in one launch.  The i-particles
// are grouped by cluster (non-decreasing cluster index); each interacts with
// its own cluster's j-particles only, with that cluster's eps2.
// Packs i-particles [i0, i1) into the pinned upload buffer: [xi | vi | (ai) |
// h2] doubles and ids (reference sapporohostclass.cpp:613-649); returns
// whether any of them has h2 > 0 (neighbour search on).
bool Engine::pack_i(int i0, int i1, const int* id, const double (*xi)[3], const double (*vi)[3],
                    const double (*ai)[3], const double* h2, double* bx, double* bv, double* ba, double* bh,
                    int* bid) {
  const bool h6 = order_ >= ORDER_H6;
  std::atomic<bool> ngb_on{false};
  par_for((size_t)(i1 - i0), [&](size_t b, size_t e) {
    b += (size_t)i0;
    e += (size_t)i0;
    const size_t n = e - b;
    memcpy(bx + 3 * b, xi[b], sizeof(double) * 3 * n);
    if (vi) memcpy(bv + 3 * b, vi[b], sizeof(double) * 3 * n);
    else memset(bv + 3 * b, 0, sizeof(double) * 3 * n);
    if (h6) {
      if (ai) memcpy(ba + 3 * b, ai[b], sizeof(double) * 3 * n);
      else memset(ba + 3 * b, 0, sizeof(double) * 3 * n);
    }
    if (h2) {
      memcpy(bh + b, h2 + b, sizeof(double) * n);
      bool any = false;
      for (size_t i = b; i < e; ++i) any |= h2[i] > 0.0;
      if (any) ngb_on.store(true, std::memory_order_relaxed);
    }  // without radii the h2 region is neither written nor uploaded: the kernels read it
       // only with the neighbour search on
    if (id) memcpy(bid + b, id + b, sizeof(int) * n);
    else std::fill(bid + b, bid + e, -1);
  });
  return ngb_on.load();
}

int Engine::start_grav_calc_clusters(int ni, const int* cluster, const int* id, const double (*xi)[3],
                                     const double (*vi)[3], const double (*ai)[3], const double* eps2c,
                                     const double* h2) {
  if (!open_) return fail("start_grav_calc_clusters: library not open");
  if (cl_off_.empty()) return fail("start_grav_calc_clusters: no cluster layout (set_clusters)");
  if (ni == 0) return 0;
  if (ni < 0 || ni > ENS_MAX_I || !cluster || !eps2c) return fail("start_grav_calc_clusters: bad arguments");
  const int nc = (int)cl_off_.size();
  // The i-particles of each active cluster: a run of equal cluster indices,
  // runs in increasing cluster order.  The run starts are found in parallel;
  // the groups are then built from them in order.  Invalid input is
  // diagnosed by the serial scan below, which names the first fault.
  std::vector<int> starts;
  {
    std::mutex mu;
    std::vector<std::pair<size_t, std::vector<int>>> found;
    std::atomic<bool> bad{false};
    par_for((size_t)ni, [&](size_t b, size_t e) {
      std::vector<int> bnd;
      bool fault = false;
      for (size_t i = b; i < e; ++i) {
        const int c = cluster[i];
        fault |= c < 0 || c >= nc;
        if (i > 0 && cluster[i - 1] != c) {
          fault |= cluster[i - 1] > c;
          bnd.push_back((int)i);
        }
      }
      if (fault) bad.store(true, std::memory_order_relaxed);
      std::lock_guard<std::mutex> lk(mu);
      found.emplace_back(b, std::move(bnd));
    });
    if (bad.load()) {
      for (int i = 0, prev = -1; i < ni; ++i) {
        const int c = cluster[i];
        if (c < 0 || c >= nc) return fail("start_grav_calc_clusters: cluster index out of range");
        if (i > 0 && c != cluster[i - 1] && c <= prev)
          return fail("start_grav_calc_clusters: i-particles not grouped by cluster");
        prev = c;
      }
    }
    std::sort(found.begin(), found.end(),
              [](const std::pair<size_t, std::vector<int>>& x, const std::pair<size_t, std::vector<int>>& y) {
                return x.first < y.first;
              });
    starts.push_back(0);
    for (const auto& f : found) starts.insert(starts.end(), f.second.begin(), f.second.end());
  }
  std::vector<EnsGroup> grps;
  std::vector<int> gcl;  // the cluster of each group
  grps.reserve(starts.size());
  gcl.reserve(starts.size());
  long work = 0;
  for (size_t q = 0; q < starts.size(); ++q) {
    const int i = starts[q], e = q + 1 < starts.size() ? starts[q + 1] : ni;
    const int c = cluster[i];
    const int n = e - i;
    const int P = n > 32 ? 64 : n > 8 ? 16 : n > 2 ? 4 : 1;
    grps.push_back(EnsGroup{i, n, P, 0, 0, 0});
    gcl.push_back(c);
    work += (long)((n + P - 1) / P) * cl_n_[c];
  }
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (ensure_icap(ni) != 0) return -1;
  bool has_h2 = false;
  if (h2)
    for (int i = 0; i < ni && !has_h2; ++i) has_h2 = h2[i] > 0.0;
  if (has_h2) {
    // packed host lists sized for a mean of 64 neighbours (larger totals fall
    // back to reading the device slots)
    if (ensure_lists(ni) != 0) return -1;
    if (ensure_packed(std::max<long>((long)NPIPES * NGB_CAP, 64L * ni)) != 0) return -1;
  }

  double* bx = ibuf_h_;
  double* bv = bx + 3 * (size_t)ni;
  double* ba = bv + 3 * (size_t)ni;
  double* bh = ba + (order_ >= ORDER_H6 ? 3 * (size_t)ni : 0);
  // Large calls pack and upload in chunks: each chunk's copies are queued as
  // soon as it is packed, so the transfer overlaps the packing of the rest.
  const bool h6 = order_ >= ORDER_H6;
  const int nchunk = ni >= (1 << 18) ? 4 : ni >= (1 << 16) ? 2 : 1;
  bool ngb_on = false;
  for (int k = 0; k < nchunk; ++k) {
    const int b = (int)((long)ni * k / nchunk), e = (int)((long)ni * (k + 1) / nchunk);
    ngb_on |= pack_i(b, e, id, xi, vi, ai, h2, bx, bv, ba, bh, iid_h_);
    const size_t n = (size_t)(e - b);
    auto up = [&](const double* h) {
      return hipMemcpyAsync(ibuf_d_ + (h - ibuf_h_), h, n * sizeof(double), hipMemcpyHostToDevice, stream_);
    };
    auto up3 = [&](const double* h) {
      return hipMemcpyAsync(ibuf_d_ + (h - ibuf_h_), h, 3 * n * sizeof(double), hipMemcpyHostToDevice, stream_);
    };
    HIPCHK(up3(bx + 3 * (size_t)b), "i upload");
    HIPCHK(up3(bv + 3 * (size_t)b), "i upload");
    if (h6) HIPCHK(up3(ba + 3 * (size_t)b), "i upload");
    if (h2) HIPCHK(up(bh + b), "i upload");
    HIPCHK(hipMemcpyAsync(iid_d_ + b, iid_h_ + b, n * sizeof(int), hipMemcpyHostToDevice, stream_), "i upload");
  }

  // work table: per active cluster, i-groups of P (by its active count) x j-splits
  const long target = (long)cus_ * 4 * WAVES_PER_SIMD_TARGET;
  const long jraw = std::max<long>(MIN_J_PER_SPLIT, (work + target - 1) / target);
  // j per split in whole 64-interaction folds of a lane: JTILE for lane shape
  // P = 64 (one j-lane per i-particle), 16 tiles for the smaller shapes (Q <= 16)
  const long jch64 = (jraw + JTILE - 1) / JTILE * JTILE;
  const long jchs = (jraw + JTILE * 16 - 1) / (JTILE * 16) * (JTILE * 16);
  // items of lane shape P = 64 first, the smaller shapes after them: each part
  // runs in a kernel of its own, so the P = 64 part keeps its register budget
  int n64 = 0, nsm = 0;
  for (size_t q = 0; q < grps.size(); ++q) {
    EnsGroup& g = grps[q];
    const long jch = g.P == 64 ? jch64 : jchs;
    g.S = std::max(1, (int)((cl_n_[gcl[q]] + jch - 1) / jch));
    (g.P == 64 ? n64 : nsm) += (g.n + g.P - 1) / g.P * g.S;
  }
  int b64 = 0, bsm = n64;
  for (EnsGroup& g : grps) {
    int& b = g.P == 64 ? b64 : bsm;
    g.first = b;
    b += (g.n + g.P - 1) / g.P * g.S;
  }
  const int nitems = n64 + nsm;
  const int ng = (int)grps.size();
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_order = al(sizeof(EnsItem) * nitems), o_grp = o_order + al(sizeof(int) * nitems);
  const size_t tab_bytes = o_grp + sizeof(EnsGroup) * ng;
  if (tab_bytes > ens_tab_cap_) {
    hfree(ens_tab_h_);
    dfree(ens_tab_d_);
    ens_tab_cap_ = std::max(tab_bytes, 2 * ens_tab_cap_);
    HIPCHK(hmalloc(&ens_tab_h_, ens_tab_cap_), "ensemble work table (host)");
    HIPCHK(dmalloc(&ens_tab_d_, ens_tab_cap_), "ensemble work table");
  }
  EnsItem* items = reinterpret_cast<EnsItem*>(ens_tab_h_);
  int* order = reinterpret_cast<int*>(ens_tab_h_ + o_order);
  for (int q = 0; q < ng; ++q) {
    const EnsGroup& g = grps[q];
    const int c = gcl[q], off = cl_off_[c], nj = cl_n_[c];
    const long jch = g.P == 64 ? jch64 : jchs;
    EnsItem* it = items + g.first;
    for (int i0 = g.ib; i0 < g.ib + g.n; i0 += g.P) {
      const int cnt = std::min(g.P, g.ib + g.n - i0);
      for (int s = 0; s < g.S; ++s) {
        const int jb = off + (int)std::min<long>((long)s * jch, nj);
        const int je = off + (int)std::min<long>((long)(s + 1) * jch, nj);
        *it++ = EnsItem{i0, cnt, jb, je, g.P, c, g.S, 0};
      }
    }
  }
  memcpy(ens_tab_h_ + o_grp, grps.data(), sizeof(EnsGroup) * ng);
  // launch order: longest j-range first within each part, equal lengths in
  // table order (a stable counting sort on the length)
  auto order_part = [&](int k0, int k1) {
    int maxlen = 0;
    for (int k = k0; k < k1; ++k) maxlen = std::max(maxlen, items[k].je - items[k].jb);
    std::vector<int> start(maxlen + 2, 0);
    for (int k = k0; k < k1; ++k) ++start[maxlen - (items[k].je - items[k].jb) + 1];
    for (int l = 1; l <= maxlen + 1; ++l) start[l] += start[l - 1];
    for (int k = k0; k < k1; ++k) order[k0 + start[maxlen - (items[k].je - items[k].jb)]++] = k;
  };
  order_part(0, n64);
  order_part(n64, nitems);
  if (ens_eps_cap_ < nc) {
    dfree(ens_eps_d_);
    ens_eps_cap_ = nc;
    HIPCHK(dmalloc(&ens_eps_d_, (size_t)nc), "ensemble softening");
  }
  HIPCHK(hipMemcpyAsync(ens_tab_d_, ens_tab_h_, tab_bytes, hipMemcpyHostToDevice, stream_), "work table upload");
  HIPCHK(hipMemcpyAsync(ens_eps_d_, eps2c, sizeof(double) * nc, hipMemcpyHostToDevice, stream_), "eps upload");
  if (launch_ens_idesc(stream_, reinterpret_cast<const EnsGroup*>(ens_tab_d_ + o_grp), ng, ni, idesc_d_) != 0)
    return fail("ensemble descriptor launch", hipGetLastError());

  int nupd = 0;
  if (upload_updates(&nupd) != 0) return -1;
  if (post_desc(ni, nupd, ngb_on ? ni : 0, 0.0) != 0) return -1;
  if (timing_) HIPCHK(hipEventRecord(ev_d0_, stream_), "event");
  if (scatter_updates(nupd, ngb_on ? ni : 0, predict_ ? cl_total_ : 0) != 0) return -1;
  if (predict_) {
    if (predict_range(0, cl_total_) != 0) return -1;
  } else {
    timed_predict_ = false;
  }
  predict_ = false;
  if (ensure_partials((size_t)nitems * 64) != 0) return -1;
  EnsArgs ea{};
  ea.rec = st_.rec;
  ea.idr = st_.idr;
  ea.ipos = ibuf_d_;
  ea.ivel = ibuf_d_ + 3 * (size_t)ni;
  ea.iacc = ibuf_d_ + 6 * (size_t)ni;
  ea.ih2 = ibuf_d_ + (bh - bx);
  ea.iid = iid_d_;
  ea.eps2 = ens_eps_d_;
  ea.items = reinterpret_cast<const EnsItem*>(ens_tab_d_);
  ea.nitems = nitems;
  ea.order = reinterpret_cast<const int*>(ens_tab_d_ + o_order);
  ea.idesc = idesc_d_;
  ea.part = part_d_;
  ea.part_ds2 = pds2_d_;
  ea.part_addr = paddr_d_;
  ea.ngb_cnt = cnt_d_;
  ea.ngb_list = list_d_;
  ea.ngb_on = ngb_on ? 1 : 0;
  FinalizeArgs fz = finalize_args(ni, ngb_on);
  fz.inorder = 0;  // ensembles: packed lists (NGB_CAP per i-particle would not fit)
  if (timing_) HIPCHK(hipEventRecord(ev_f0_, stream_), "event");
  if (launch_force_ens(stream_, order_, f64_, ea, fz, n64) != 0)
    return fail("ensemble force launch", hipGetLastError());
  if (timing_) {
    HIPCHK(hipEventRecord(ev_f1_, stream_), "event");
    HIPCHK(hipEventRecord(ev_d1_, stream_), "event");
  }
  timed_force_ = timing_;
  HIPCHK(hipEventRecord(ev_done_, stream_), "event");
  done_rec_seq_ = ++rec_seq_;
  inflight_ = true;
  have_results_ = true;
  ngb_fetched_ = false;
  lists_in_slots_ = false;
  lists_packed_ = true;
  ngb_on_last_ = ngb_on;
  ni_last_ = ni;
  nj_last_ = cl_total_;
  last_cfg_ = ForceConfig{0, nitems, (int)jch64, 0, 0};
  jhost_valid_ = false;
  return 0;
}

// Reference startGravCalc (sapporohostclass.cpp:586-703).
int Engine::start_grav_calc(int nj, int ni, const int* id, const double (*xi)[3], const double (*vi)[3],
                            const double (*ai)[3], double eps2, const double* h2) {
  if (!open_) return fail("start_grav_calc: library not open");
  if (ni == 0 || nj == 0) return 0;  // reference 597
  if (ni < 0 || nj < 0) return fail("start_grav_calc: negative count");
  if (ni > NPIPES) return fail("start_grav_calc: ni exceeds n_pipes");  // reference assert, 604
  if (ensure_icap(NPIPES) != 0) return -1;
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (ensure_store(nj) != 0) return -1;

  // pack i-particles: [xi | vi | ai | h2] doubles and ids (reference 613-649)
  // into the call block, once the copy that last read it has completed
  if (desc_pending_ && desc_seq_ > done_seq_) HIPCHK(hipEventSynchronize(ev_desc_), "call block wait");
  desc_pending_ = false;
  int bucket = 1;
  while (bucket < ni) bucket <<= 1;
  int* bid = reinterpret_cast<int*>(reinterpret_cast<char*>(call_h_) + CALL_HDR);
  double* bx = reinterpret_cast<double*>(reinterpret_cast<char*>(call_h_) + CALL_HDR + call_ids_bytes(bucket));
  call_iid_d_ = reinterpret_cast<int*>(reinterpret_cast<char*>(call_d_) + CALL_HDR);
  call_ibuf_d_ = reinterpret_cast<double*>(reinterpret_cast<char*>(call_d_) + CALL_HDR + call_ids_bytes(bucket));
  double* bv = bx + 3 * (size_t)ni;
  double* ba = bv + 3 * (size_t)ni;
  double* bh = ba + (order_ >= ORDER_H6 ? 3 * (size_t)ni : 0);
  const bool ngb_on = pack_i(0, ni, id, xi, vi, ai, h2, bx, bv, ba, bh, bid);
  const size_t nd = (size_t)((h2 ? bh + ni : bh) - bx);
  if (ngb_on) {
    if (ensure_lists(NPIPES) != 0) return -1;
    if (ensure_packed((long)NPIPES * NGB_CAP) != 0) return -1;
  }

  int nupd = 0;
  if (upload_updates(&nupd) != 0) return -1;
  // descriptor, ids and doubles in one copy
  desc_h_->ti = t_i_;
  desc_h_->eps2 = eps2;
  desc_h_->ni = ni;
  desc_h_->nupd = nupd;
  desc_h_->ni_reset = ngb_on ? ni : 0;
  desc_h_->pad = 0;
  const size_t nbytes = (size_t)((char*)(bx + nd) - (char*)call_h_);
  HIPCHK(hipMemcpyAsync(call_d_, call_h_, nbytes, hipMemcpyHostToDevice, stream_), "call upload");
  HIPCHK(hipEventRecord(ev_desc_, stream_), "event");
  desc_pending_ = true;
  desc_seq_ = ++rec_seq_;
  const bool predict = predict_;
  predict_ = false;  // reference 698
  const ForceConfig cfg = choose_config(ni, nj, ngb_on);
  // With graphs the grid covers the power-of-two bucket of ni (calls of one
  // bucket share a captured sequence); direct launches cover ni exactly.  The
  // finalize shape follows the bucket in both cases, so results agree.
  const int grid_ni = graphs_ ? std::min(bucket, NPIPES) : ni;
  if (ensure_partials((size_t)cfg.S * grid_ni) != 0) return -1;
  const int sblocks = scatter_blocks(nupd, ngb_on ? ni : 0);
  // device window: everything after the host->device copies
  if (timing_) HIPCHK(hipEventRecord(ev_d0_, stream_), "event");
  if (graphs_) {
    if (launch_call_graph(grid_ni, nj, cfg, ngb_on, predict, sblocks, bucket >= 1024) != 0) return -1;
    timed_predict_ = timed_force_ = false;
  } else {
    if (enqueue_call(grid_ni, nj, cfg, ngb_on, predict, sblocks, bucket >= 1024) != 0) return -1;
    if (!predict) timed_predict_ = false;
    timed_force_ = timing_;
  }
  if (timing_) HIPCHK(hipEventRecord(ev_d1_, stream_), "event");
  HIPCHK(hipEventRecord(ev_done_, stream_), "event");
  done_rec_seq_ = ++rec_seq_;
  inflight_ = true;
  have_results_ = true;
  ngb_fetched_ = false;
  lists_in_slots_ = false;
  lists_packed_ = false;
  ngb_on_last_ = ngb_on;
  ni_last_ = ni;
  nj_last_ = nj;
  last_cfg_ = cfg;
  jhost_valid_ = false;
  return 0;
}

// The device work of one force call: [scatter] -> [predict] -> force ->
// finalize.  Every per-call value comes from the descriptor, so the sequence
// depends only on its arguments here (the key of a captured graph).
int Engine::enqueue_call(int grid_ni, int nj, const ForceConfig& cfg, bool ngb_on, bool predict, int sblocks,
                         bool rows) {
  if (sblocks > 0) {
    ScatterArgs a{};
    a.st = st_;
    a.upd = upd_d_;
      a.nf = raw_fields(order_);
    a.ngb_cnt = cnt_d_;
    a.pack_ctr = pack_ctr_d_;
    a.cd = desc_d_;
    a.grid_blocks = sblocks;
    a.pred_j1 = predict ? nj : 0;
    const bool ev = timing_ && !capturing_;
    if (ev) HIPCHK(hipEventRecord(ev_s0_, stream_), "event");
    if (launch_scatter(stream_, order_, f64_, a) != 0) return fail("scatter launch", hipGetLastError());
    if (ev) HIPCHK(hipEventRecord(ev_s1_, stream_), "event");
    timed_scatter_ = ev;
    last_nupd_ = desc_h_->nupd;
  } else {
    timed_scatter_ = false;
  }
  if (predict && predict_range(0, nj) != 0) return -1;
  ForceArgs fa{};
  fa.rec = st_.rec;
  fa.idr = st_.idr;
  fa.nj = nj;
  fa.ibuf = call_ibuf_d_;
  fa.iid = call_iid_d_;
  fa.cd = desc_d_;
  fa.grid_ni = grid_ni;
  fa.part = part_d_;
  fa.part_ds2 = pds2_d_;
  fa.part_addr = paddr_d_;
  fa.ngb_cnt = cnt_d_;
  fa.ngb_list = list_d_;
  fa.ngb_on = ngb_on ? 1 : 0;
  const bool ev = timing_ && !capturing_;
  if (ev) HIPCHK(hipEventRecord(ev_f0_, stream_), "event");
  if (launch_force(stream_, order_, f64_, fa, cfg) != 0) return fail("force launch", hipGetLastError());
  if (ev) HIPCHK(hipEventRecord(ev_f1_, stream_), "event");
  FinalizeArgs fz = finalize_args(grid_ni, ngb_on);
  fz.S = partial_rows(cfg);
  fz.rows = rows ? 1 : 0;
  if (launch_finalize(stream_, order_, f64_, fz) != 0) return fail("finalize launch", hipGetLastError());
  return 0;
}

// The allocation-free call: same launch shape as start_grav_calc (the j-split
// is coarsened only when its partial rows would not fit the scratch region),
// same kernels, everything queued on s.  Preconditions instead of the
// allocating steps of a regular call: no staged j-updates and nj within the
// store (flush with force_prediction first).
int Engine::grav_calc_scratch(void* scratch, size_t bytes, int nj, int ni, const int* id_d, const double* ibuf_d,
                              double eps2, double* out_d, hipStream_t s) {
  if (!open_) return fail("grav_calc_noalloc: library not open");
  if (ni <= 0 || nj <= 0 || ni > NPIPES) return fail("grav_calc_noalloc: need 0 < ni <= 16384 and nj > 0");
  if (!scratch || !id_d || !ibuf_d || !out_d) return fail("grav_calc_noalloc: null pointer");
  if (nj > st_.cap) return fail("grav_calc_noalloc: nj beyond the j-store (set and flush the j-particles first)");
  if (nstaged_ > 0) return fail("grav_calc_noalloc: staged j-updates pending (flush with force_prediction first)");
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (!s) s = stream_;
  const int NF = part_fields(order_);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  ForceConfig cfg = choose_config(ni, nj, false);
  auto need = [&](const ForceConfig& c) {
    const size_t r = (size_t)partial_rows(c) * ni;
    return al(sizeof(CallDesc)) + al(sizeof(double) * NF * r) + al(sizeof(double) * r) + al(sizeof(int) * r);
  };
  while (need(cfg) > bytes && cfg.S > 1) {
    cfg.jchunk *= 2;
    cfg.S = (nj + cfg.jchunk - 1) / cfg.jchunk;
  }
  if (need(cfg) > bytes) return fail("grav_calc_noalloc: call too large for the scratch region");
  const size_t r = (size_t)partial_rows(cfg) * ni;
  char* p = static_cast<char*>(scratch);
  CallDesc* cd = reinterpret_cast<CallDesc*>(p);
  p += al(sizeof(CallDesc));
  double* part = reinterpret_cast<double*>(p);
  p += al(sizeof(double) * NF * r);
  double* pds2 = reinterpret_cast<double*>(p);
  p += al(sizeof(double) * r);
  int* paddr = reinterpret_cast<int*>(p);
  const CallDesc v{t_i_, eps2, ni, 0, 0, 0};
  if (launch_write_desc(s, cd, v) != 0) return fail("descriptor launch", hipGetLastError());
  if (predict_) {
    PredictArgs pa{};
    pa.st = st_;
    pa.j0 = 0;
    pa.j1 = nj;
    pa.cd = cd;
    if (launch_predict(s, order_, f64_, pa) != 0) return fail("predict launch", hipGetLastError());
    predict_ = false;
    jhost_valid_ = false;
  }
  ForceArgs fa{};
  fa.rec = st_.rec;
  fa.idr = st_.idr;
  fa.nj = nj;
  fa.ibuf = ibuf_d;
  fa.iid = id_d;
  fa.cd = cd;
  fa.grid_ni = ni;
  fa.part = part;
  fa.part_ds2 = pds2;
  fa.part_addr = paddr;
  fa.ngb_on = 0;
  if (launch_force(s, order_, f64_, fa, cfg) != 0) return fail("force launch", hipGetLastError());
  int bucket = 1;
  while (bucket < ni) bucket <<= 1;
  FinalizeArgs fz{};
  fz.part = part;
  fz.part_ds2 = pds2;
  fz.part_addr = paddr;
  fz.S = partial_rows(cfg);
  fz.cd = cd;
  fz.grid_ni = ni;
  fz.rows = bucket >= 1024 ? 1 : 0;
  fz.rec = st_.rec;
  fz.out = out_d;
  fz.ostride = ni;
  fz.ngb_on = 0;
  if (launch_finalize(s, order_, f64_, fz) != 0) return fail("finalize launch", hipGetLastError());
  return 0;
}

constexpr size_t MAX_GRAPHS = 64;

// Replays the captured sequence of this launch shape, capturing it first if
// needed (least recently used graphs are dropped beyond MAX_GRAPHS; graphs
// whose buffers were reallocated are dropped when next seen).
int Engine::launch_call_graph(int grid_ni, int nj, const ForceConfig& cfg, bool ngb_on, bool predict, int sblocks,
                              bool rows) {
  const GraphKey key{grid_ni, nj, cfg.P, cfg.S, cfg.jchunk, cfg.smajor, cfg.pipe, ngb_on ? 1 : 0,
                     predict ? 1 : 0, sblocks * 2 + (rows ? 1 : 0), gen_};
  for (auto& e : graph_cache_) {
    if (e.key == key) {
      e.used = ++tick_;
      HIPCHK(hipGraphLaunch(e.exec, stream_), "graph launch");
      return 0;
    }
  }
  bool stale = false;
  for (auto& e : graph_cache_) stale |= e.key.gen != gen_;
  if (stale || graph_cache_.size() >= MAX_GRAPHS) {
    HIPCHK(hipStreamSynchronize(stream_), "graph cache sync");
    for (size_t k = 0; k < graph_cache_.size();) {
      if (graph_cache_[k].key.gen != gen_) {
        (void)hipGraphExecDestroy(graph_cache_[k].exec);
        graph_cache_.erase(graph_cache_.begin() + (long)k);
      } else {
        ++k;
      }
    }
    if (graph_cache_.size() >= MAX_GRAPHS) {
      size_t lru = 0;
      for (size_t k = 1; k < graph_cache_.size(); ++k)
        if (graph_cache_[k].used < graph_cache_[lru].used) lru = k;
      (void)hipGraphExecDestroy(graph_cache_[lru].exec);
      graph_cache_.erase(graph_cache_.begin() + (long)lru);
    }
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "graph capture");
  capturing_ = true;
  const int rc = enqueue_call(grid_ni, nj, cfg, ngb_on, predict, sblocks, rows);
  capturing_ = false;
  const hipError_t ec = hipStreamEndCapture(stream_, &g);
  if (rc != 0 || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc != 0 ? rc : fail("graph capture", ec);
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail("graph instantiate", ei);
  graph_cache_.push_back(GraphEntry{key, x, ++tick_});
  HIPCHK(hipGraphLaunch(x, stream_), "graph launch");
  return 0;
}

void Engine::drop_graphs() {
  if (graph_cache_.empty()) return;
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : graph_cache_) (void)hipGraphExecDestroy(e.exec);
  graph_cache_.clear();
}

// Waits for the last call's completion event by polling it (the caller is
// blocked anyway; polling wakes within a microsecond or two, where a blocking
// wait costs tens).  Everything recorded before that event is then complete.
int Engine::wait_results() {
  if (!inflight_) return 0;
  hipError_t e;
  while ((e = hipEventQuery(ev_done_)) == hipErrorNotReady) {
    for (int k = 0; k < 32; ++k) _mm_pause();
  }
  if (e != hipSuccess) return fail("result wait", e);
  HIPCHK(hipGetLastError(), "device error");
  inflight_ = false;
  done_seq_ = done_rec_seq_;
  return 0;
}

float Engine::last_force_ms() {
  if (!timing_ || !have_results_ || !timed_force_) return 0.f;
  (void)hipEventSynchronize(ev_f1_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev_f0_, ev_f1_) != hipSuccess) return 0.f;
  return ms;
}

float Engine::last_device_ms() {
  if (!timing_ || !have_results_) return 0.f;
  (void)hipEventSynchronize(ev_d1_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev_d0_, ev_d1_) != hipSuccess) return 0.f;
  return ms;
}

float Engine::last_finalize_ms() {
  if (!timing_ || !have_results_ || !timed_force_) return 0.f;
  (void)hipEventSynchronize(ev_d1_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev_f1_, ev_d1_) != hipSuccess) return 0.f;
  return ms;
}

float Engine::last_scatter_ms(int* nupd) {
  if (nupd) *nupd = timed_scatter_ ? last_nupd_ : 0;
  if (!timing_ || !timed_scatter_) return 0.f;
  (void)hipEventSynchronize(ev_s1_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev_s0_, ev_s1_) != hipSuccess) return 0.f;
  return ms;
}

float Engine::last_predict_ms() {
  if (!timing_ || !timed_predict_) return 0.f;
  (void)hipEventSynchronize(ev_p1_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev_p0_, ev_p1_) != hipSuccess) return 0.f;
  return ms;
}

// Reference getGravResults (sapporohostclass.cpp:705-802): outputs are
// written for [0, ni); the nearest neighbour is reported only when its ds2
// (float) is below 1e10, otherwise nnb = 0 and dsmin is left untouched.
int Engine::get_grav_results(int nj, int ni, double (*acc)[3], double (*jrk)[3], double (*snp)[3],
                             double (*crk)[3], double* pot, int* nnb, double* dsmin, bool ngb) {
  if (!open_) return fail("get_grav_results: library not open");
  if (ni == 0 || nj == 0) return 0;  // reference 718
  if (!have_results_ || ni > ni_last_) return fail("get_grav_results: no matching firsthalf");
  if (wait_results() != 0) return -1;
  const int NF = part_fields(order_);
  // result rows are field-major in the host-mapped block
  const bool h6 = order_ >= ORDER_H6;
  par_for((size_t)ni, [&](size_t b, size_t e) {
    auto vec3 = [&](double (*dst)[3], int f0) {
      if (!dst) return;
      const double *x = result_field(f0), *y = result_field(f0 + 1), *z = result_field(f0 + 2);
      for (size_t i = b; i < e; ++i) { dst[i][0] = x[i]; dst[i][1] = y[i]; dst[i][2] = z[i]; }
    };
    vec3(acc, 0);
    if (pot) memcpy(pot + b, result_field(3) + b, sizeof(double) * (e - b));
    vec3(jrk, 4);
    if (h6) {
      vec3(snp, 7);
      if (crk)
        for (size_t i = b; i < e; ++i) crk[i][0] = crk[i][1] = crk[i][2] = 0.0;
    }
    if (ngb) {
      const double *ds = result_field(NF), *nn = result_field(NF + 1);
      for (size_t i = b; i < e; ++i) {
        if (nnb) nnb[i] = 0;
        if (ds[i] < 1.0e10) {
          if (nnb) nnb[i] = (int)nn[i];
          if (dsmin) dsmin[i] = ds[i];
        }
      }
    }
  });
  return 0;
}

// Reference read_ngb_list (sapporohostclass.cpp:804-835): the lists of the
// last call; returns 1 if any i-particle has >= 256 neighbours.  The finalize
// wrote them sorted into host-mapped memory (in i order within regions of
// LIST_REGION i-particles, or, for ensembles, packed by work-group), so once
// the call is complete nothing is copied, unless an
// ensemble's packed total exceeded its buffer: then the sorted device slots
// are copied back instead.
int Engine::read_ngb_list() {
  if (!open_ || !have_results_) return 0;
  if (wait_results() != 0) return -1;
  const int ni = ni_last_;
  const double* cnt = result_field(part_fields(order_) + 2);
  int overflow = 0;
  long total = 0;
  for (int i = 0; i < ni; ++i) {
    const int c = (int)cnt[i];
    overflow |= c >= NGB_CAP;
    total += std::min(c, NGB_CAP);
  }
  if (ngb_on_last_ && lists_packed_ && total > packed_cap_ && !lists_in_slots_) {
    slot_lists_h_.resize((size_t)ni * NGB_CAP);
    HIPCHK(hipMemcpyAsync(slot_lists_h_.data(), list_d_, sizeof(int) * (size_t)ni * NGB_CAP, hipMemcpyDeviceToHost,
                          stream_), "list slot download");
    HIPCHK(hipStreamSynchronize(stream_), "list slot sync");
    lists_in_slots_ = true;
  }
  ngb_fetched_ = true;
  return overflow;
}

// Reference get_ngb_list (sapporohostclass.cpp:837-880): copies up to
// maxlength ids (sorted ascending), n = the neighbour count, returns 1 when
// the count reaches maxlength.
int Engine::get_ngb_list(int ipipe, int maxlen, int* n, int* list) {
  if (n) *n = 0;
  if (!open_ || !have_results_) return 0;
  if (!ngb_fetched_) {
    const int r = read_ngb_list();
    if (r < 0) return r;
  }
  if (ipipe < 0 || ipipe >= ni_last_) return fail("get_ngb_list: ipipe out of range");
  const int len = ngb_on_last_ ? (int)result_field(part_fields(order_) + 2)[ipipe] : 0;
  const int c = std::min(len, NGB_CAP);
  const int ncopy = std::max(0, std::min(c, maxlen));
  if (list && ncopy > 0) {
    const int* src = lists_in_slots_ ? slot_lists_h_.data() + (size_t)ipipe * NGB_CAP : packed_h_ + loff_h_[ipipe];
    memcpy(list, src, sizeof(int) * (size_t)ncopy);
  }
  if (n) *n = len;
  const bool overflow = len >= maxlen;
  if (overflow)
    fprintf(stderr, "sapporo2_amd: get_ngb_list overflow for ipipe= %d, ngb= %d\n", ipipe, len);
  return overflow ? 1 : 0;
}

// Reference forcePrediction (sapporohostclass.cpp:1164-1195).
int Engine::force_prediction(int nj) {
  if (!open_) return fail("force_prediction: library not open");
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (flush_updates(0, predict_ ? std::max(nj, 0) : 0) != 0) return -1;
  if (predict_ && nj > 0) {
    if (predict_range(0, nj) != 0) return -1;
  }
  predict_ = false;
  jhost_valid_ = false;
  return 0;
}

// Read-back of the j-store (predicted and raw state) for [0, max address].
int Engine::fetch_j_host(int need) {
  if (jhost_valid_ && need <= jhost_n_) return 0;
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (flush_updates(0) != 0) return -1;
  const int n = std::max(need, max_addr_ + 1);
  if (n <= 0) return 0;
  if (ensure_store(n) != 0) return -1;
  if (gather_cap_ < (size_t)n) {
    dfree(gather_d_);
    HIPCHK(dmalloc(&gather_d_, (size_t)n * 10), "gather buffer");
    gather_cap_ = n;
  }
  GatherArgs g{};
  g.st = st_;
  g.n = n;
  g.out = gather_d_;
  g.addr = nullptr;
  if (launch_gather_pred(stream_, order_, f64_, g) != 0) return fail("gather launch", hipGetLastError());
  const int nfr = raw_fields(order_);
  jpred_h_.resize((size_t)n * 10);
  jraw_h_.resize((size_t)n * nfr);
  jid_h_.resize(n);
  HIPCHK(hipMemcpyAsync(jpred_h_.data(), gather_d_, sizeof(double) * (size_t)n * 10, hipMemcpyDeviceToHost,
                        stream_), "gather download");
  for (int f = 0; f < nfr; ++f)
    HIPCHK(hipMemcpyAsync(jraw_h_.data() + (size_t)f * n, st_.raw + (size_t)f * st_.cap, sizeof(double) * n,
                          hipMemcpyDeviceToHost, stream_), "raw download");
  HIPCHK(hipMemcpyAsync(jid_h_.data(), st_.id, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, stream_),
         "id download");
  HIPCHK(hipStreamSynchronize(stream_), "read-back sync");
  jhost_valid_ = true;
  jhost_n_ = n;
  return 0;
}

// Reference retrieve_predicted_j_particle (sapporohostclass.cpp:969-1041).
int Engine::retrieve_predicted_j_particles(int n, const int* addr, double* mass, double (*pos)[3],
                                           double (*vel)[3], double (*acc)[3]) {
  if (!open_) return fail("retrieve_predicted_j_particles: library not open");
  if (n < 0 || (n > 0 && !addr)) return fail("retrieve_predicted_j_particles: bad arguments");
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (flush_updates(0) != 0) return -1;
  for (int k = 0; k < n; ++k)
    if (addr[k] < 0 || addr[k] > max_addr_) return fail("retrieve_predicted_j_particles: address never set");
  if (gather_cap_ < (size_t)n) {
    dfree(gather_d_);
    HIPCHK(dmalloc(&gather_d_, (size_t)n * 10), "gather buffer");
    gather_cap_ = n;
  }
  if (gaddr_cap_ < (size_t)n) {
    dfree(gaddr_d_);
    HIPCHK(dmalloc(&gaddr_d_, (size_t)n), "gather addresses");
    gaddr_cap_ = n;
  }
  HIPCHK(hipMemcpyAsync(gaddr_d_, addr, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, stream_), "address upload");
  GatherArgs g{};
  g.st = st_;
  g.n = n;
  g.out = gather_d_;
  g.addr = gaddr_d_;
  if (launch_gather_pred(stream_, order_, f64_, g) != 0) return fail("gather launch", hipGetLastError());
  gather_h_.resize((size_t)n * 10);
  HIPCHK(hipMemcpyAsync(gather_h_.data(), gather_d_, sizeof(double) * (size_t)n * 10, hipMemcpyDeviceToHost,
                        stream_), "gather download");
  HIPCHK(hipStreamSynchronize(stream_), "read-back sync");
  for (int k = 0; k < n; ++k) {
    const double* p = gather_h_.data() + (size_t)k * 10;
    if (pos) { pos[k][0] = p[0]; pos[k][1] = p[1]; pos[k][2] = p[2]; }
    if (mass) mass[k] = p[3];
    if (vel) { vel[k][0] = p[4]; vel[k][1] = p[5]; vel[k][2] = p[6]; }
    if (acc) { acc[k][0] = p[7]; acc[k][1] = p[8]; acc[k][2] = p[9]; }
  }
  return 0;
}

int Engine::retrieve_predicted_j_particle(int addr, double* mass, double* id, double* eps2, double pos[3],
                                          double vel[3], double acc[3]) {
  if (!open_) return fail("retrieve_predicted_j_particle: library not open");
  if (addr < 0) return fail("retrieve_predicted_j_particle: negative address");
  if (fetch_j_host(addr + 1) != 0) return -1;
  if (addr >= jhost_n_) return fail("retrieve_predicted_j_particle: address never set");
  const double* p = jpred_h_.data() + (size_t)addr * 10;
  if (pos) { pos[0] = p[0]; pos[1] = p[1]; pos[2] = p[2]; }
  if (mass) *mass = p[3];
  if (vel) { vel[0] = p[4]; vel[1] = p[5]; vel[2] = p[6]; }
  if (eps2) *eps2 = jraw_h_[(size_t)RF_EPS * jhost_n_ + addr];
  if (acc && order_ >= ORDER_H6) { acc[0] = p[7]; acc[1] = p[8]; acc[2] = p[9]; }
  if (id) *id = (double)jid_h_[addr];
  return 0;
}

// Checkpoint layout: StateHeader, cluster sizes (ncl ints), then for the n
// stored addresses: raw fields (nfr x n doubles, field-major), ids (n ints),
// predicted records (n x record bytes), [ppos, pvel, (pacc): 3 x n doubles
// each, mixed precision], tile id ranges (ceil(n / JTILE) int2).  Sections
// start at 8-byte offsets.
namespace {
constexpr uint64_t STATE_MAGIC = 0x31534d4150415341ull;  // "ASAPAMS1"
struct StateHeader {
  uint64_t magic;
  int version, order, f64, n, max_addr, nj_max, ncl, predict;
  double t_i;
};
size_t pad8(size_t b) { return (b + 7) / 8 * 8; }
}  // namespace

long long Engine::state_bytes() {
  if (!open_) return fail("state_bytes: library not open");
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (flush_updates(0) != 0) return -1;
  const size_t n = (size_t)std::max(max_addr_ + 1, cl_total_);
  const size_t nfr = raw_fields(order_);
  size_t b = pad8(sizeof(StateHeader)) + pad8(sizeof(int) * cl_n_.size());
  b += sizeof(double) * nfr * n + pad8(sizeof(int) * n) + pad8(jrec_bytes(order_, f64_) * n);
  if (!f64_) b += sizeof(double) * 3 * n * (order_ >= ORDER_H6 ? 3 : 2);
  b += sizeof(int2) * ((n + JTILE - 1) / JTILE);
  return (long long)b;
}

int Engine::save_state(void* buf, long long bytes) {
  const long long need = state_bytes();
  if (need < 0) return -1;
  if (!buf || bytes < need) return fail("save_state: buffer smaller than state_bytes()");
  const int n = std::max(max_addr_ + 1, cl_total_);
  if (ensure_store(n) != 0) return -1;
  char* p = static_cast<char*>(buf);
  StateHeader h{};
  h.magic = STATE_MAGIC; h.version = 1; h.order = order_; h.f64 = f64_ ? 1 : 0; h.n = n;
  h.max_addr = max_addr_; h.nj_max = nj_max_; h.ncl = (int)cl_n_.size(); h.predict = predict_ ? 1 : 0;
  h.t_i = t_i_;
  memcpy(p, &h, sizeof(h));
  p += pad8(sizeof(h));
  if (!cl_n_.empty()) memcpy(p, cl_n_.data(), sizeof(int) * cl_n_.size());
  p += pad8(sizeof(int) * cl_n_.size());
  auto d2h = [&](const void* src, size_t b) -> int {
    if (b) HIPCHK(hipMemcpyAsync(p, src, b, hipMemcpyDeviceToHost, stream_), "state download");
    p += pad8(b);
    return 0;
  };
  const size_t cap = st_.cap;
  for (int f = 0; f < raw_fields(order_); ++f)
    if (d2h(st_.raw + f * cap, sizeof(double) * n) != 0) return -1;
  if (d2h(st_.id, sizeof(int) * n) != 0) return -1;
  if (d2h(st_.rec, jrec_bytes(order_, f64_) * n) != 0) return -1;
  for (double* a : {st_.ppos, st_.pvel, st_.pacc})
    if (a)
      for (int d = 0; d < 3; ++d)
        if (d2h(a + d * cap, sizeof(double) * n) != 0) return -1;
  if (d2h(st_.idr, sizeof(int2) * ((n + JTILE - 1) / JTILE)) != 0) return -1;
  HIPCHK(hipStreamSynchronize(stream_), "state download sync");
  return 0;
}

int Engine::load_state(const void* buf, long long bytes) {
  if (!open_) return fail("load_state: library not open");
  if (!buf || bytes < (long long)sizeof(StateHeader)) return fail("load_state: buffer too small");
  StateHeader h;
  memcpy(&h, buf, sizeof(h));
  if (h.magic != STATE_MAGIC || h.version != 1) return fail("load_state: not a state dump");
  if (h.order != order_ || h.f64 != (f64_ ? 1 : 0))
    return fail("load_state: dump was made with another order or precision");
  if (h.n < 0 || h.ncl < 0 || h.max_addr >= h.n) return fail("load_state: corrupt header");
  HIPCHK(hipSetDevice(device_), "hipSetDevice");
  if (inflight_ && wait_results() != 0) return -1;
  if (flush_updates(0) != 0) return -1;  // staged updates are superseded by the dump
  const char* p = static_cast<const char*>(buf) + pad8(sizeof(h));
  const size_t n = h.n;
  const size_t nfr = raw_fields(order_);
  size_t need = pad8(sizeof(h)) + pad8(sizeof(int) * (size_t)h.ncl) + sizeof(double) * nfr * n +
                pad8(sizeof(int) * n) + pad8(jrec_bytes(order_, f64_) * n) +
                (f64_ ? 0 : sizeof(double) * 3 * n * (order_ >= ORDER_H6 ? 3 : 2)) +
                sizeof(int2) * ((n + JTILE - 1) / JTILE);
  if ((size_t)bytes < need) return fail("load_state: buffer shorter than the dump it holds");
  std::vector<int> sizes((size_t)h.ncl);
  if (h.ncl) memcpy(sizes.data(), p, sizeof(int) * (size_t)h.ncl);
  p += pad8(sizeof(int) * (size_t)h.ncl);
  cl_off_.clear(); cl_n_.clear(); cl_total_ = 0;
  if (h.ncl && set_clusters(h.ncl, sizes.data()) != 0) return -1;
  if (cl_total_ > h.n) return fail("load_state: corrupt cluster sizes");
  nj_max_ = std::max({h.nj_max, h.max_addr + 1, cl_total_});  // the source's logical capacity
  if ((int)slot_of_.size() < nj_max_) slot_of_.resize(nj_max_, -1);
  if (ensure_store(std::max(h.n, 1)) != 0) return -1;
  // clear the whole store: addresses beyond the dump's were never set there
  const size_t cap = st_.cap;
  HIPCHK(hipMemsetAsync(st_.raw, 0, sizeof(double) * nfr * cap, stream_), "memset");
  HIPCHK(hipMemsetAsync(st_.id, 0, sizeof(int) * cap, stream_), "memset");
  HIPCHK(hipMemsetAsync(st_.rec, 0, jrec_bytes(order_, f64_) * cap, stream_), "memset");
  HIPCHK(hipMemsetAsync(st_.idr, 0, sizeof(int2) * (cap / JTILE), stream_), "memset");
  for (double* a : {st_.ppos, st_.pvel, st_.pacc})
    if (a) HIPCHK(hipMemsetAsync(a, 0, sizeof(double) * 3 * cap, stream_), "memset");
  auto h2d = [&](void* dst, size_t b) -> int {
    if (b) HIPCHK(hipMemcpyAsync(dst, p, b, hipMemcpyHostToDevice, stream_), "state upload");
    p += pad8(b);
    return 0;
  };
  for (size_t f = 0; f < nfr; ++f)
    if (h2d(st_.raw + f * cap, sizeof(double) * n) != 0) return -1;
  if (h2d(st_.id, sizeof(int) * n) != 0) return -1;
  if (h2d(st_.rec, jrec_bytes(order_, f64_) * n) != 0) return -1;
  for (double* a : {st_.ppos, st_.pvel, st_.pacc})
    if (a)
      for (int d = 0; d < 3; ++d)
        if (h2d(a + d * cap, sizeof(double) * n) != 0) return -1;
  if (h2d(st_.idr, sizeof(int2) * ((n + JTILE - 1) / JTILE)) != 0) return -1;
  HIPCHK(hipStreamSynchronize(stream_), "state upload sync");
  max_addr_ = h.max_addr;
  t_i_ = h.t_i;
  predict_ = h.predict != 0;
  jhost_valid_ = false;
  have_results_ = false;
  ngb_fetched_ = false;
  return 0;
}

// Reference retrieve_j_particle_state (sapporohostclass.cpp:1045-1147).
int Engine::retrieve_j_particle_state(int addr, double* mass, double* id, double* eps2, double pos[3],
                                      double vel[3], double acc[3], double jrk[3], double ppos[3],
                                      double pvel[3], double pacc[3]) {
  if (!open_) return fail("retrieve_j_particle_state: library not open");
  if (addr < 0) return fail("retrieve_j_particle_state: negative address");
  if (fetch_j_host(addr + 1) != 0) return -1;
  if (addr >= jhost_n_) return fail("retrieve_j_particle_state: address never set");
  const size_t n = jhost_n_;
  auto raw = [&](int f) { return jraw_h_[(size_t)f * n + addr]; };
  const double* p = jpred_h_.data() + (size_t)addr * 10;
  if (mass) *mass = raw(RF_M);
  if (pos) for (int d = 0; d < 3; ++d) pos[d] = raw(RF_X + d);
  if (ppos) for (int d = 0; d < 3; ++d) ppos[d] = p[d];
  if (pvel) for (int d = 0; d < 3; ++d) pvel[d] = p[4 + d];
  if (eps2) *eps2 = raw(RF_EPS);
  if (vel) for (int d = 0; d < 3; ++d) vel[d] = raw(RF_VX + d);
  if (acc) for (int d = 0; d < 3; ++d) acc[d] = raw(RF_AX + d);
  if (jrk) for (int d = 0; d < 3; ++d) jrk[d] = raw(RF_JX + d);
  if (id) *id = (double)jid_h_[addr];
  if (pacc && order_ >= ORDER_H6) for (int d = 0; d < 3; ++d) pacc[d] = p[7 + d];
  return 0;
}

}  // namespace sap
