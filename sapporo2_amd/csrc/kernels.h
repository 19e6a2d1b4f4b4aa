// kernels.h -- interface between the host engine and the HIP kernels.
//
// The engine (engine.cpp) is the only caller; it owns every device buffer and
// passes plain pointers.  All kernels are written for gfx950 (wave64, 256 CUs).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sapk {

// Integration order / working precision, numbered as the reference's enums
// (lib/include/defines.h:30-31).
enum Order { ORDER_G5 = 0, ORDER_H4 = 1, ORDER_H6 = 2 };
enum Precision { PREC_FLOAT = 0, PREC_DS = 1, PREC_DOUBLE = 2 };

// Per-i neighbour slots: the reference's NGB_PB (defines.h:9).
constexpr int NGB_CAP = 256;

// Raw j-store: one double array per field ("field-major SoA"), each of the
// store's capacity.  Fields follow the reference's set_j_particle payload
// (sapporohostclass.cpp:500-516): time, dt, mass, eps, pos, vel, acc, jerk,
// snap, crackle.
enum RawField {
  RF_T = 0, RF_DT, RF_M, RF_EPS,
  RF_X, RF_Y, RF_Z, RF_VX, RF_VY, RF_VZ,
  RF_AX, RF_AY, RF_AZ, RF_JX, RF_JY, RF_JZ,
  RF_SX, RF_SY, RF_SZ, RF_CX, RF_CY, RF_CZ,
  RF_COUNT
};
// Number of raw fields an update record carries for an order.
constexpr int raw_fields(int order) { return order >= ORDER_H6 ? RF_COUNT : RF_SX; }

// Double-single coordinate: hi = x rounded to float, lo = (x - hi) rounded
// (the reference's DSX, kernels.cu:154-178).
struct DSF { float hi, lo; };

// Predicted j-records read by the force kernels (one array of structs, read
// wave-uniformly through the scalar cache or per lane).  Mixed precision keeps
// positions as double-single pairs; relative positions are
// (xj.hi - xi.hi) + (xj.lo - xi.lo) in FP32.  FP64 keeps FP64 positions.
struct JRecM4 { DSF x, y, z; float m, vx, vy, vz; int id, pad; };                  // 48 B
struct JRecM6 { DSF x, y, z; float m, vx, vy, vz, ax, ay, az; int id, pad[2]; };   // 64 B
struct JRecD4 { double x, y, z, m, vx, vy, vz; int id, pad; };                        // 64 B
struct JRecD6 { double x, y, z, m, vx, vy, vz, ax, ay, az; int id, pad[3]; };         // 96 B

inline size_t jrec_bytes(int order, bool f64) {
  if (order >= ORDER_H6) return f64 ? sizeof(JRecD6) : sizeof(JRecM6);
  return f64 ? sizeof(JRecD4) : sizeof(JRecM4);
}

// Number of doubles per i-particle result record written by the finalize
// kernel: acc[3], pot, jerk[3], [snap[3]], dsmin, nnb (as double), count.
constexpr int out_fields(int order) { return order >= ORDER_H6 ? 13 : 10; }
// Number of FP64 partial-force fields: acc[3], pot, jerk[3], [snap[3]].
constexpr int part_fields(int order) { return order >= ORDER_H6 ? 10 : 7; }

// ---------------------------------------------------------------------------
// Addresses are grouped in tiles of JTILE; per tile the store keeps a
// conservative [min id, max id] range so the force kernel can skip the
// per-pair self-interaction test on tiles that cannot hold an i-particle's id.
constexpr int JTILE = 64;

struct StoreView {
  double* raw;       // RF_COUNT (or raw_fields) x cap doubles
  int* id;           // cap ints
  void* rec;         // cap predicted records
  double* ppos;      // mixed precision only: 3 x cap FP64 predicted position
  double* pvel;      // mixed precision only: 3 x cap FP64 predicted velocity
  double* pacc;      // mixed precision SIXTH only: 3 x cap FP64 predicted acc
  int2* idr;         // cap / JTILE tile id ranges (x = min, y = max)
  int cap;           // multiple of JTILE
};

// Per-call values the kernels read from device memory, so that a launch
// sequence does not depend on them and can be captured once and replayed.
struct CallDesc {
  double ti;             // block time of the predictor
  double eps2;           // softening of the force call
  int ni;                // active i-particles of the call
  int nupd;              // staged j-updates to scatter
  int ni_reset;          // neighbour counters to clear before the call
  int pad;
};

// Scatter staged updates into the store and reset the per-i neighbour
// counters of the coming force call (fused: one launch, grid-stride over
// cd->nupd and cd->ni_reset).
struct ScatterArgs {
  StoreView st;
  const double* upd;     // nupd rows of nf + 1 doubles: fields 0..nf-1 of RawField, then (address, id) as an int2
  int nf;
  int* ngb_cnt;          // zeroed for [0, ni_reset)
  int* pack_ctr;         // neighbour-list pack counter, zeroed when ni_reset > 0 (may be null)
  const CallDesc* cd;
  int grid_blocks;       // launch size (host)
  int pred_j1;           // a prediction of [0, pred_j1) follows in the same call: the predicted
                         // records of updated addresses below it are left to the predictor
};

struct PredictArgs {
  StoreView st;
  int j0, j1;            // predict addresses [j0, j1)
  const CallDesc* cd;    // ti
};

struct ForceArgs {
  const void* rec;       // predicted records
  const int2* idr;       // tile id ranges (StoreView::idr)
  int nj;
  const double* ibuf;    // [xi (ni x 3) | vi (ni x 3) | ai (ni x 3, SIXTH) | h2 (ni)]
  const int* iid;        // ni
  const CallDesc* cd;    // ni, eps2
  int grid_ni;           // launch size (host): i-particles the grid covers, >= cd->ni
  double* part;          // part_fields x S x ni FP64 partial sums
  double* part_ds2;      // S x ni  nearest-neighbour ds2 (working precision value)
  int* part_addr;        // S x ni  nearest-neighbour address (INT_MAX: none)
  int* ngb_cnt;          // ni exact neighbour counts (atomic)
  int* ngb_list;         // ni x NGB_CAP ids (arbitrary order until finalized)
  int ngb_on;            // neighbour search enabled (some h2 > 0)
};

// Launch geometry of one force call: each wave owns P i-particles and Q = 64/P
// j-lanes per i-particle over one of S contiguous j-ranges of jchunk.
// smajor: wave-to-split mapping; pipe: neighbour-id stores deferred by one group.
struct ForceConfig { int P; int S; int jchunk; int smajor; int pipe; };

// Partial-sum rows a force launch writes per i-particle.  With S a multiple of
// 4 and split-minor wave order the 4 waves of a work-group are 4 consecutive
// splits of one i-group; they combine their sums in LDS and write one row.
inline int partial_rows(const ForceConfig& c) { return (c.S % 4 == 0 && !c.smajor) ? c.S / 4 : c.S; }

// Results and neighbour lists go straight to host-mapped pinned memory (no
// copy command): field-major result rows out[f * ostride + i], and each sorted
// list at packed[list_off[i] ...].  Single-system calls (inorder = 1) pack the
// lists in i order within regions of LIST_REGION i-particles, region r from
// packed[r * LIST_REGION * NGB_CAP], so the host reads them as one stream; an
// offset is its region's base plus the stored lengths before it there.
// Ensembles (inorder = 0, more i-particles than a region layout would fit)
// take offsets from pack_ctr, one atomic per work-group.  A packed list that
// would end beyond packed_cap is not written to the host; every sorted list is
// also kept in ngb_list (device).
constexpr int LIST_REGION = 256;
struct FinalizeArgs {
  const double* part; const double* part_ds2; const int* part_addr;
  int S;
  const CallDesc* cd;    // ni
  int grid_ni;           // launch size (host), >= cd->ni
  int rows;              // shape: 1 thread per i-particle (else a work-group per i-particle)
  const void* rec;       // to look up the neighbour's id
  int* ngb_cnt; int* ngb_list;
  double* out;           // host-mapped: out_fields x ostride doubles, field-major
  long ostride;
  int* list_off;         // host-mapped: ni list offsets into packed
  int* packed;           // host-mapped: sorted ids of all lists, packed_cap ints
  long packed_cap;
  int* pack_ctr;         // device: next free slot of packed (zeroed by the scatter)
  int inorder;           // 1: i-ordered regions (single system); 0: atomic offsets (ensembles)
  int ngb_on;
};

// Ensemble work item: i-particles [i0, i0 + icount) of one cluster (icount <=
// P, P in {64, 16, 4, 1}) against that cluster's j in [jb, je), one of the
// nsplit items that cover the cluster's j for these i-particles.
struct EnsItem { int i0, icount, jb, je, P, cluster, nsplit, pad; };

// The i-particles [ib, ib + n) of one cluster in an ensemble call: lane shape P,
// S splits per i-group, items first .. first + ceil(n / P) * S - 1.
struct EnsGroup { int ib, n, P, S, first, pad; };

struct EnsArgs {
  const void* rec; const int2* idr;
  const double* ipos; const double* ivel; const double* iacc; const double* ih2; const int* iid;
  const double* eps2;        // per cluster
  const EnsItem* items; int nitems;
  const int* order;          // item indices, longest first within each part
  const int4* idesc;         // per i: first item, number of splits, i-slot
  double* part; double* part_ds2; int* part_addr;   // part_fields x nitems x 64, nitems x 64
  int* ngb_cnt; int* ngb_list; int ngb_on;
};

// Gathers the FP64 predicted state (pos, mass, vel, [acc]) into a dense AoS
// buffer for read-back, 10 doubles per row: of all j in [0, n) (addr null) or
// of the n listed addresses.
struct GatherArgs { StoreView st; int n; double* out; const int* addr; };

int launch_scatter(hipStream_t s, int order, bool f64, const ScatterArgs& a);
int launch_predict(hipStream_t s, int order, bool f64, const PredictArgs& a);
int launch_force(hipStream_t s, int order, bool f64, const ForceArgs& a, const ForceConfig& c);
int launch_finalize(hipStream_t s, int order, bool f64, const FinalizeArgs& a);
// Ensemble force + finalize (+ list sort) over a work table whose first n64
// items have lane shape P = 64; one single-wave work-group per item.  Without
// neighbour lists, an item that covers all of its cluster's j (nsplit = 1)
// writes its i-particles' result rows itself, and the finalize skips them.
int launch_force_ens(hipStream_t s, int order, bool f64, const EnsArgs& a, const FinalizeArgs& f, int n64);
int launch_gather_pred(hipStream_t s, int order, bool f64, const GatherArgs& a);
// Per i-particle of an ensemble call (ni of them, in ng groups): {first item of
// its i-group, split count, slot in the i-group, 1 for the small-shape part}.
int launch_ens_idesc(hipStream_t s, const EnsGroup* g, int ng, int ni, int4* idesc);
// Writes the call descriptor v to device memory cd with a one-thread kernel
// (no copy command, nothing allocated).
int launch_write_desc(hipStream_t s, CallDesc* cd, const CallDesc& v);

// The P values with compiled force kernels.
bool force_p_supported(int P);

}  // namespace sapk
