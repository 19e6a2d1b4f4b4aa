// noalloc.hip -- sapamd_grav_calc_noalloc: a force call that allocates nothing.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the call descriptor and the partial sums of
// the force kernels live in `scratch`, a file-scope device array of
// SCRATCH_BYTES that each GPU the library runs on holds until the process
// ends.  Every call on one GPU shares that GPU's array: calls that overlap on
// one GPU (two streams, two threads) race.  Nothing in it needs clearing (every
// partial row and the descriptor are written before they are read).
#include <mutex>

#include "../../include/sapporo_amd.h"
#include "engine.h"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;   // 64 MiB
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// Device address of `scratch` on each device, looked up once per device with
// that device current (each GPU holds its own copy at its own address).  The
// array's name is a host-side stand-in: only these addresses reach kernels.
void* g_scratch[MAX_DEVICES];
std::mutex g_scratch_mu;

int scratch_address(int dev, void** ptr) {
  if (dev < 0 || dev >= MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return -1;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return -1;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return -1;
    g_scratch[dev] = p;
  }
  *ptr = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" {

long long sapamd_noalloc_scratch_bytes(void) { return (long long)SCRATCH_BYTES; }

int sapamd_grav_calc_noalloc(sapamd_engine* e, int nj, int ni, const int* id, const double* ibuf, double eps2,
                             double* out, void* stream) {
  if (!e) return -1;
  void* p = nullptr;
  if (scratch_address(e->eng.device(), &p) != 0) {
    fprintf(stderr, "sapporo2_amd: grav_calc_noalloc: scratch address lookup failed\n");
    return -1;
  }
  return e->eng.grav_calc_scratch(p, SCRATCH_BYTES, nj, ni, id, ibuf, eps2, out,
                                  reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
