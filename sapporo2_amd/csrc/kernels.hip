// kernels.hip -- gfx950 kernels of the direct-summation engine.
//
//   k_scatter     staged j-updates -> FP64 SoA store + predicted records
//                 (reference dev_copy_particles, kernels.cu:245-301), fused with
//                 the per-call reset of neighbour counters (dev_reset_buffers,
//                 kernels.cu:1109-1144).
//   k_predict     FP64 Horner predictor, HBM-bound (dev_predictor, kernels.cu:307-411).
//   k_force       acc / pot / jerk [/ snap] + nearest neighbour + neighbour list
//                 (dev_evaluate_gravity_*, kernels.cu:552-1041).  One wave owns P
//                 i-particles held in registers and Q = 64/P j-lanes per
//                 i-particle over one contiguous j-range (the "j-split").  With
//                 P = 64 every lane of a wave reads the same j, so the j-record
//                 is fetched once per wave through the scalar cache into SGPRs
//                 and feeds the VALU directly.  Mixed precision: double-single
//                 positions, relative positions and pair terms in FP32,
//                 per-tile FP32 sums folded into FP64 accumulators.
//   k_finalize    deterministic fixed-order reduction of the S partial sums,
//                 nearest-neighbour selection, neighbour-list sort and packing;
//                 results and packed lists are written straight into
//                 host-mapped pinned memory (no copy command per call).
//
// Built with -ffp-contract=off: every fused multiply-add below is explicit.
#include "kernels.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

namespace sapk {

template <int ORDER, bool F64> struct RecOf;
template <> struct RecOf<ORDER_H4, false> { using T = JRecM4; };
template <> struct RecOf<ORDER_H6, false> { using T = JRecM6; };
template <> struct RecOf<ORDER_H4, true> { using T = JRecD4; };
template <> struct RecOf<ORDER_H6, true> { using T = JRecD6; };

static_assert(sizeof(JRecM4) == 48, "record layout");
static_assert(sizeof(JRecM6) == 64, "record layout");
static_assert(sizeof(JRecD4) == 64, "record layout");
static_assert(sizeof(JRecD6) == 96, "record layout");

// Double-single split (reference DSX::operator=(double), kernels.cu:163-167).
__device__ __forceinline__ DSF to_ds(double x) {
  DSF d;
  d.hi = (float)x;
  d.lo = (float)(x - (double)d.hi);
  return d;
}
__device__ __forceinline__ void put_pos(double& dst, double x) { dst = x; }
__device__ __forceinline__ void put_pos(DSF& dst, double x) { dst = to_ds(x); }

// Write the working-precision record of particle j from FP64 values.
template <int ORDER, bool F64>
__device__ __forceinline__ void put_rec(void* base, int j, double x, double y, double z, double m,
                                        double vx, double vy, double vz, double ax, double ay,
                                        double az, int id) {
  using Rec = typename RecOf<ORDER, F64>::T;
  using R = typename std::conditional<F64, double, float>::type;
  Rec r{};
  put_pos(r.x, x); put_pos(r.y, y); put_pos(r.z, z);
  r.m = (R)m; r.vx = (R)vx; r.vy = (R)vy; r.vz = (R)vz;
  if constexpr (ORDER >= ORDER_H6) { r.ax = (R)ax; r.ay = (R)ay; r.az = (R)az; }
  r.id = id;
  reinterpret_cast<Rec*>(base)[j] = r;
}

// ---------------------------------------------------------------------------
// Scatter of staged updates (+ neighbour counter reset).
template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int ni_reset = a.cd->ni_reset, nupd = a.cd->nupd;
  for (int k = tid; k < ni_reset; k += stride) a.ngb_cnt[k] = 0;
  if (tid == 0 && ni_reset > 0 && a.pack_ctr) *a.pack_ctr = 0;
  const size_t cap = (size_t)a.st.cap;
  for (int k = tid; k < nupd; k += stride) {
    const double* u = a.upd + (size_t)k * (a.nf + 1);
    const int2 pair = *reinterpret_cast<const int2*>(u + a.nf);  // address, id
    const int addr = pair.x;
    const int id = pair.y;
    if (addr < 0 || addr >= a.st.cap) continue;  // host guarantees range
    for (int f = 0; f < a.nf; ++f) a.st.raw[(size_t)f * cap + addr] = u[f];
    a.st.id[addr] = id;
    // widen the tile's id range (conservative until the next prediction recomputes it)
    atomicMin(&a.st.idr[addr / JTILE].x, id);
    atomicMax(&a.st.idr[addr / JTILE].y, id);
    // predicted = raw for updated particles (kernels.cu:281-298), unless the
    // predictor that follows rewrites this address anyway
    if (addr < a.pred_j1) continue;
    double ax = 0, ay = 0, az = 0;
    if constexpr (ORDER >= ORDER_H6) { ax = u[RF_AX]; ay = u[RF_AY]; az = u[RF_AZ]; }
    put_rec<ORDER, F64>(a.st.rec, addr, u[RF_X], u[RF_Y], u[RF_Z], u[RF_M], u[RF_VX], u[RF_VY],
                        u[RF_VZ], ax, ay, az, id);
    if constexpr (!F64) {
      a.st.ppos[addr] = u[RF_X]; a.st.ppos[cap + addr] = u[RF_Y]; a.st.ppos[2 * cap + addr] = u[RF_Z];
      a.st.pvel[addr] = u[RF_VX]; a.st.pvel[cap + addr] = u[RF_VY]; a.st.pvel[2 * cap + addr] = u[RF_VZ];
      if constexpr (ORDER >= ORDER_H6) {
        a.st.pacc[addr] = ax; a.st.pacc[cap + addr] = ay; a.st.pacc[2 * cap + addr] = az;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 predictor.  The expression is the reference's (kernels.cu:332-383),
// evaluated without contraction, including the zero terms for FOURTH order.
template <int ORDER>
__device__ __forceinline__ void predict_one(const double* __restrict__ R, size_t cap, int j,
                                            double ti, double p[3], double v[3], double ac[3]) {
  const double dt = ti - R[RF_T * cap + j];
  const double dt2 = (1. / 2.) * dt;
  const double dt3 = (1. / 3.) * dt;
  double dt4 = (1. / 4.) * dt;
  double dt5 = (1. / 5.) * dt;
  if constexpr (ORDER <= ORDER_H4) { dt4 = 0; dt5 = 0; }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xs = R[(RF_X + d) * cap + j];
    const double vs = R[(RF_VX + d) * cap + j];
    const double as = R[(RF_AX + d) * cap + j];
    const double js = R[(RF_JX + d) * cap + j];
    double ss = 0.0, cs = 0.0;
    if constexpr (ORDER >= ORDER_H6) { ss = R[(RF_SX + d) * cap + j]; cs = R[(RF_CX + d) * cap + j]; }
    p[d] = xs + dt * (vs + dt2 * (as + dt3 * (js + dt4 * (ss + dt5 * (cs)))));
    v[d] = vs + dt * (as + dt2 * (js + dt3 * (ss + dt4 * (cs))));
    ac[d] = as;
    if constexpr (ORDER >= ORDER_H6) ac[d] = as + dt * (js + dt2 * (ss + dt3 * (cs)));
  }
}

template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_predict(PredictArgs a) {
  const int j = a.j0 + blockIdx.x * blockDim.x + threadIdx.x;
  // exact id range of this wave's tile (j0 is a multiple of JTILE = wave size);
  // addresses past j1 are included so later calls with a larger nj stay covered
  {
    const int id = j < a.st.cap ? a.st.id[j] : 0;
    int lo = j < a.st.cap ? id : INT_MAX, hi = j < a.st.cap ? id : INT_MIN;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      lo = min(lo, __shfl_xor(lo, m));
      hi = max(hi, __shfl_xor(hi, m));
    }
    if ((threadIdx.x & 63) == 0 && j < a.st.cap) a.st.idr[j / JTILE] = make_int2(lo, hi);
  }
  if (j >= a.j1) return;
  const size_t cap = (size_t)a.st.cap;
  double p[3], v[3], ac[3];
  predict_one<ORDER>(a.st.raw, cap, j, a.cd->ti, p, v, ac);
  put_rec<ORDER, F64>(a.st.rec, j, p[0], p[1], p[2], a.st.raw[RF_M * cap + j], v[0], v[1], v[2],
                      ac[0], ac[1], ac[2], a.st.id[j]);
  if constexpr (!F64) {
    a.st.ppos[j] = p[0]; a.st.ppos[cap + j] = p[1]; a.st.ppos[2 * cap + j] = p[2];
    a.st.pvel[j] = v[0]; a.st.pvel[cap + j] = v[1]; a.st.pvel[2 * cap + j] = v[2];
    if constexpr (ORDER >= ORDER_H6) {
      a.st.pacc[j] = ac[0]; a.st.pacc[cap + j] = ac[1]; a.st.pacc[2 * cap + j] = ac[2];
    }
  }
}

// ---------------------------------------------------------------------------
// Force kernel.
template <typename R> __device__ __forceinline__ R fmaR(R a, R b, R c);
template <> __device__ __forceinline__ float fmaR<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmaR<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// 1/sqrt: v_rsq_f32 (1 ulp) in FP32; in FP64 v_rsq_f64 plus one cubic
// correction step, full double precision for the positive normal operands
// used here (r2 >= eps2 > 0; other cases are masked by the caller).
template <typename R> __device__ __forceinline__ R rsq(R x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double rsq<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// Relative coordinate xj - xi: FP64 subtraction, or the double-single
// difference (xj.hi - xi.hi) + (xj.lo - xi.lo) in FP32 (kernels.cu:170-178).
__device__ __forceinline__ double to_rel(double xj, double xi) { return xj - xi; }
__device__ __forceinline__ float to_rel(const DSF& xj, const DSF& xi) { return (xj.hi - xi.hi) + (xj.lo - xi.lo); }

template <typename R> struct PosOf { using T = DSF; };
template <> struct PosOf<double> { using T = double; };
__device__ __forceinline__ void get_pos(double& dst, double x) { dst = x; }
__device__ __forceinline__ void get_pos(DSF& dst, double x) { dst = to_ds(x); }

// Per-lane state of one i-particle.
template <int ORDER, typename R>
struct IState {
  typename PosOf<R>::T x, y, z;
  R vx, vy, vz, ax, ay, az, h2;
  int id, i;
};

template <typename R> __device__ __forceinline__ R minR(R a, R b);
template <> __device__ __forceinline__ float minR<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> __device__ __forceinline__ double minR<double>(double a, double b) { return __builtin_fmin(a, b); }

// Relative position and squared distance of one pair, the arithmetic every
// distance test uses (the nearest-neighbour resolution recomputes it and must
// agree bit for bit).
template <typename R, typename Rec, typename St>
__device__ __forceinline__ R pair_ds2(const Rec& r, const St& s, R& dx, R& dy, R& dz) {
  dx = to_rel(r.x, s.x);
  dy = to_rel(r.y, s.y);
  dz = to_rel(r.z, s.z);
  return fmaR<R>(dz, dz, fmaR<R>(dy, dy, dx * dx));
}

constexpr int NF4 = 7, NF6 = 10;
template <int ORDER> constexpr int nf() { return ORDER >= ORDER_H6 ? NF6 : NF4; }

// One pair interaction (kernels.cu:565-667) accumulated into t[].  Nearest
// neighbour: with BLOCKMIN, dsmin receives the pair's squared distance (+inf
// for an excluded pair) for the caller's group minimum; otherwise dsmin and
// nnaddr hold the distance and the address j of the nearest so far.
template <int ORDER, bool NGB, bool CHECK, bool BLOCKMIN, typename R, typename Rec>
__device__ __forceinline__ bool pair(const Rec& r, const int j, const bool ok, const IState<ORDER, R>& s,
                                     const R eps2, R (&t)[nf<ORDER>()], R& dsmin, int& nnaddr,
                                     int* __restrict__ ngb_cnt, int* __restrict__ ngb_list) {
  R dx, dy, dz;
  const R ds2 = pair_ds2<R>(r, s, dx, dy, dz);
  // bitwise & keeps the predicates branch-free (no exec-masked regions)
  const bool use = CHECK ? (ok & (r.id != s.id)) : ok;
  const R inv0 = rsq<R>(ds2 + eps2);
  const R inv = use ? inv0 : R(0);
  const R invr2 = inv * inv;
  const R mr1 = r.m * inv;
  const R mr3 = mr1 * invr2;
  t[0] = fmaR<R>(mr3, dx, t[0]);
  t[1] = fmaR<R>(mr3, dy, t[1]);
  t[2] = fmaR<R>(mr3, dz, t[2]);
  t[3] = t[3] - mr1;
  const R dvx = r.vx - s.vx;
  const R dvy = r.vy - s.vy;
  const R dvz = r.vz - s.vz;
  const R rv = fmaR<R>(dz, dvz, fmaR<R>(dy, dvy, dx * dvx));
  if constexpr (ORDER == ORDER_H4) {
    // jerk += m/r^3 dv - 3 (m/r^3)(dr.dv)/r^2 dr   (kernels.cu:633-636)
    const R q = (rv * invr2) * R(-3);
    t[4] = fmaR<R>(mr3, fmaR<R>(q, dx, dvx), t[4]);
    t[5] = fmaR<R>(mr3, fmaR<R>(q, dy, dvy), t[5]);
    t[6] = fmaR<R>(mr3, fmaR<R>(q, dz, dvz), t[6]);
  } else {
    // kernels.cu:642-667: alpha = dr.dv/r^2, beta = (dv^2 + dr.da)/r^2 + alpha^2
    const R dax = r.ax - s.ax;
    const R day = r.ay - s.ay;
    const R daz = r.az - s.az;
    const R v2 = fmaR<R>(dvz, dvz, fmaR<R>(dvy, dvy, dvx * dvx));
    const R ra = fmaR<R>(dz, daz, fmaR<R>(dy, day, dx * dax));
    const R alpha = rv * invr2;
    const R beta = fmaR<R>(v2 + ra, invr2, alpha * alpha);
    const R m3a = alpha * R(-3);
    // jerk_ij = m/r^3 (dv - 3 alpha dr)
    const R jx = mr3 * fmaR<R>(m3a, dx, dvx);
    const R jy = mr3 * fmaR<R>(m3a, dy, dvy);
    const R jz = mr3 * fmaR<R>(m3a, dz, dvz);
    // snap_ij = m/r^3 (da - 3 beta dr) - 6 alpha jerk_ij
    const R m3b = beta * R(-3);
    const R m6a = alpha * R(-6);
    t[7] = fmaR<R>(m6a, jx, fmaR<R>(mr3, fmaR<R>(m3b, dx, dax), t[7]));
    t[8] = fmaR<R>(m6a, jy, fmaR<R>(mr3, fmaR<R>(m3b, dy, day), t[8]));
    t[9] = fmaR<R>(m6a, jz, fmaR<R>(mr3, fmaR<R>(m3b, dz, daz), t[9]));
    t[4] = t[4] + jx;
    t[5] = t[5] + jy;
    t[6] = t[6] + jz;
  }
  // nearest neighbour: strict <, first address wins inside a lane (kernels.cu:539-546)
  if constexpr (BLOCKMIN) {
    dsmin = CHECK ? (use ? ds2 : R(INFINITY)) : ds2;
  } else {
    const bool better = use & (ds2 < dsmin);
    dsmin = better ? ds2 : dsmin;
    nnaddr = better ? j : nnaddr;
  }
  // neighbour predicate h2_i > ds2 (kernels.cu:584); the caller appends
  return NGB & use & (s.h2 > ds2);
}

// Append id to the i-particle's list; rare (~32 per 1e5 pairs), so a global
// atomic takes the slot; the count stays exact past NGB_CAP.
__device__ __forceinline__ void ngb_append(int* __restrict__ ngb_cnt, int* __restrict__ ngb_list, int i, int id) {
  const int k = atomicAdd(&ngb_cnt[i], 1);
  if (k < NGB_CAP) ngb_list[(size_t)i * NGB_CAP + k] = id;
}

// Neighbour slots reserved by one group whose id stores are still pending.
struct PendingNgb {
  int k[4], id[4];
  bool m[4];
};

// Store the ids of the previous group (their slot atomics were issued one
// group earlier, so the atomic latency overlaps this group's arithmetic).
__device__ __forceinline__ void ngb_flush(PendingNgb& p, int* __restrict__ ngb_list, int i) {
  if (p.m[0] | p.m[1] | p.m[2] | p.m[3]) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (p.m[u] && p.k[u] < NGB_CAP) ngb_list[(size_t)i * NGB_CAP + p.k[u]] = p.id[u];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) p.m[u] = false;
}

// Empty register constraint: the values are computed before this point.
template <int ORDER, typename R>
__device__ __forceinline__ void settle(R (&t)[nf<ORDER>()]) {
#pragma unroll
  for (int f = 0; f < nf<ORDER>(); ++f) asm volatile("" : "+v"(t[f]));
}

// Four consecutive j of one lane with a single neighbour branch for the group.
// The group's arithmetic is settled before the branch; otherwise the compiler
// sinks it below the (rarely taken) append block and keeps the four relative
// positions live across it, which costs a wave of occupancy.
template <int ORDER, bool NGB, bool CHECK, int Q, bool PIPE, typename R, typename Rec>
__device__ __forceinline__ void pair4(const Rec* __restrict__ rec, const int j0, const IState<ORDER, R>& s,
                                      const R eps2, R (&t)[nf<ORDER>()], R& smin, int* __restrict__ ngb_cnt,
                                      int* __restrict__ ngb_list, PendingNgb& pend) {
  R d[4];
  int idj[4];
  int unused = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = j0 + u * Q;
    const Rec& r = rec[j];
    idj[u] = r.id;
    pair<ORDER, NGB, CHECK, true>(r, j, true, s, eps2, t, d[u], unused, ngb_cnt, ngb_list);
  }
  // the group's nearest distance: it lowers the sub-block minimum, and h2 > it
  // is the test whether any of the four pairs is a neighbour
  const R g = minR<R>(minR<R>(minR<R>(d[0], d[1]), d[2]), d[3]);
  smin = minR<R>(smin, g);
  if constexpr (NGB) settle<ORDER>(t);
  // without lists, keep the scheduler from hoisting the scalar loads of a whole
  // unrolled sub-block (they exhaust the SGPRs and spill)
  if constexpr (!NGB) __builtin_amdgcn_sched_barrier(0);
  const bool any = NGB && __ballot(s.h2 > g) != 0;
  if constexpr (NGB && PIPE) {
    ngb_flush(pend, ngb_list, s.i);
    if (any) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool nb = s.h2 > d[u];
        if (nb) pend.k[u] = atomicAdd(&ngb_cnt[s.i], 1);
        pend.id[u] = idj[u];
        pend.m[u] = nb;
      }
    }
  } else if constexpr (NGB) {
    if (any) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (s.h2 > d[u]) ngb_append(ngb_cnt, ngb_list, s.i, idj[u]);
    }
  }
}

constexpr int KT = 64;  // interactions per lane between FP64 folds

// The work of one wave: i-particles [i0, i0 + icount) (icount <= P) against
// j in [jb, je), Q = 64/P j-lanes per i-particle.  On return the lanes with
// writer set hold the wave-reduced FP64 sums T, nearest-neighbour (ds2,
// address) and their i-slot.
template <int ORDER, bool F64, int P, bool NGB, bool PIPE = false>
__device__ __forceinline__ void wave_force(
    const typename RecOf<ORDER, F64>::T* __restrict__ rec, const int2* __restrict__ idr,
    const double* __restrict__ ipos, const double* __restrict__ ivel, const double* __restrict__ iacc,
    const double* __restrict__ ih2, const int* __restrict__ iid, const int i0, const int icount,
    const int jb, const int je, const double eps2d, int* __restrict__ ngb_cnt, int* __restrict__ ngb_list,
    double (&T)[nf<ORDER>()], double& dsd, int& nnaddr, int& islot, bool& writer) {
  using R = typename std::conditional<F64, double, float>::type;
  using Rec = typename RecOf<ORDER, F64>::T;
  constexpr int Q = 64 / P;
  constexpr int NF = nf<ORDER>();
  const int lane = threadIdx.x & 63;
  islot = (Q == 1) ? lane : (lane % P);
  const int jsub = (Q == 1) ? 0 : (lane / P);
  const bool ivalid = islot < icount;
  const int il = i0 + (ivalid ? islot : 0);

  IState<ORDER, R> st;
  get_pos(st.x, ipos[3 * il]); get_pos(st.y, ipos[3 * il + 1]); get_pos(st.z, ipos[3 * il + 2]);
  st.vx = (R)ivel[3 * il]; st.vy = (R)ivel[3 * il + 1]; st.vz = (R)ivel[3 * il + 2];
  st.ax = st.ay = st.az = R(0);
  if constexpr (ORDER >= ORDER_H6) {
    st.ax = (R)iacc[3 * il]; st.ay = (R)iacc[3 * il + 1]; st.az = (R)iacc[3 * il + 2];
  }
  st.h2 = (NGB && ivalid) ? (R)ih2[il] : R(-1);
  st.id = iid[il];
  st.i = il;
  const R eps2 = (R)eps2d;

#pragma unroll
  for (int f = 0; f < NF; ++f) T[f] = 0.0;
  R dsmin = R(INFINITY);
  // nearest pair so far: -1 none; >= 0 the first address of the sub-block of
  // 16 pairs (whole tiles) that holds it; <= -2 the exact address a, as -2 - a
  // (the partial tile at the end of the range)
  int nnblk = -1;
  PendingNgb pend;
#pragma unroll
  for (int u = 0; u < 4; ++u) { pend.m[u] = false; pend.k[u] = 0; pend.id[u] = 0; }

  const int nk = (je > jb) ? (je - jb + Q - 1) / Q : 0;  // iterations of the lane with jsub 0
  // FP32 pair terms are summed per fold of KT and folded into FP64 T; in FP64
  // the pair terms go straight into t, which becomes T at the end
  R t[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) t[f] = R(0);
  for (int k0 = 0; k0 < nk; k0 += KT) {
    if constexpr (!F64) {
#pragma unroll
      for (int f = 0; f < NF; ++f) t[f] = R(0);
    }
    const int kn = min(KT, nk - k0);
    const int jt = jb + k0 * Q + jsub;
    if (kn == KT && jb + (k0 + KT) * Q <= je) {
      // P = 64: one JTILE-aligned tile per fold; skip the self test when no
      // lane's id can occur in the tile
      bool check = true;
      if constexpr (Q == 1) {
        if ((jt & (JTILE - 1)) == 0) {
          const int2 r = idr[jt / JTILE];
          check = __ballot(st.id >= r.x && st.id <= r.y) != 0;
        }
      }
      // Nearest neighbour: each pair only lowers the minimum of its sub-block of
      // 16 pairs (one v_min); a sub-block that beats the lane's nearest so far
      // records its first address, resolved to the pair after the loop.
      if (check) {
        for (int k = 0; k < KT; k += 16) {
          R sm = R(INFINITY);
#pragma unroll
          for (int q = 0; q < 16; q += 4)
            pair4<ORDER, NGB, true, Q, PIPE>(rec, jt + (k + q) * Q, st, eps2, t, sm, ngb_cnt, ngb_list, pend);
          const bool better = sm < dsmin;
          dsmin = better ? sm : dsmin;
          nnblk = better ? jt + k * Q : nnblk;
        }
      } else {
        for (int k = 0; k < KT; k += 16) {
          R sm = R(INFINITY);
#pragma unroll
          for (int q = 0; q < 16; q += 4)
            pair4<ORDER, NGB, false, Q, PIPE>(rec, jt + (k + q) * Q, st, eps2, t, sm, ngb_cnt, ngb_list, pend);
          const bool better = sm < dsmin;
          dsmin = better ? sm : dsmin;
          nnblk = better ? jt + k * Q : nnblk;
        }
      }
    } else {
      if constexpr (NGB && PIPE) ngb_flush(pend, ngb_list, st.i);
      const R before = dsmin;
      int cand = 0;
      for (int k = 0; k < kn; ++k) {
        const int j = jt + k * Q;
        const bool ok = j < je;
        const Rec& r = rec[ok ? j : jb];
        if (pair<ORDER, NGB, true, false>(r, j, ok, st, eps2, t, dsmin, cand, ngb_cnt, ngb_list))
          ngb_append(ngb_cnt, ngb_list, st.i, r.id);
      }
      if (dsmin < before) nnblk = -2 - cand;  // a later pair, strictly nearer: its address is exact
    }
    if constexpr (!F64) {
#pragma unroll
      for (int f = 0; f < NF; ++f) T[f] += (double)t[f];
    }
  }
  if constexpr (F64) {
#pragma unroll
    for (int f = 0; f < NF; ++f) T[f] = (double)t[f];
  }
  // The nearest pair is the first of its sub-block at distance dsmin: a
  // sub-block replaced the candidate only when strictly nearer, so this is the
  // lane's first address at the minimum, as the per-pair rule gives.
  nnaddr = nnblk <= -2 ? -2 - nnblk : INT_MAX;
  if (nnblk >= 0) {
    bool found = false;
    for (int u = 0; u < 16; ++u) {
      const int j = nnblk + u * Q;
      const Rec& r = rec[j];
      R dx, dy, dz;
      const R d2 = pair_ds2<R>(r, st, dx, dy, dz);
      const bool hit = !found && (r.id != st.id) && d2 == dsmin;
      nnaddr = hit ? j : nnaddr;
      found = found || hit;
    }
  }

  if constexpr (NGB && PIPE) ngb_flush(pend, ngb_list, st.i);
  dsd = (double)dsmin;
  if constexpr (Q > 1) {
#pragma unroll
    for (int m = P; m < 64; m <<= 1) {
#pragma unroll
      for (int f = 0; f < NF; ++f) T[f] += __shfl_xor(T[f], m);
      const double od = __shfl_xor(dsd, m);
      const int oa = __shfl_xor(nnaddr, m);
      const bool take = (od < dsd) || (od == dsd && oa < nnaddr);
      dsd = take ? od : dsd;
      nnaddr = take ? oa : nnaddr;
    }
  }
  writer = ivalid && jsub == 0;
}

template <int ORDER, bool F64, int P, bool NGB, bool PIPE>
__global__ __launch_bounds__(256) void k_force(
    const typename RecOf<ORDER, F64>::T* __restrict__ rec, const int2* __restrict__ idr, const int nj,
    const double* __restrict__ ibuf, const int* __restrict__ iid, const CallDesc* __restrict__ cd,
    double* __restrict__ part, double* __restrict__ part_ds2, int* __restrict__ part_addr,
    int* __restrict__ ngb_cnt, int* __restrict__ ngb_list, const int S, const int jchunk, const int smajor,
    const int blockred) {
  constexpr int NF = nf<ORDER>();
  // the call's values come from the descriptor; the grid may cover more
  // i-groups than there are (a launch shape shared by a range of ni)
  const int ni = cd->ni;
  const double eps2d = cd->eps2;
  const double* ipos = ibuf;
  const double* ivel = ibuf + 3 * (size_t)ni;
  const double* iacc = ibuf + 6 * (size_t)ni;
  const double* ih2 = ibuf + (ORDER >= ORDER_H6 ? 9 : 6) * (size_t)ni;
  // wave-uniform by construction; readfirstlane lets the compiler keep the
  // j-loop bookkeeping and (for P = 64) the j-record loads in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int G = (ni + P - 1) / P;
  if (wave >= G * S) return;  // whole wave (a whole block when S % 4 == 0)
  // smajor: consecutive waves (a block's 4) share one j-range and so its
  // scalar-cache lines; otherwise consecutive waves walk the splits of one group
  const int g = smajor ? wave % G : wave / S;
  const int sidx = smajor ? wave / G : wave - g * S;
  const int jb = sidx * jchunk;
  const int je = min(nj, jb + jchunk);
  double T[NF];
  double dsd;
  int nnaddr, islot;
  bool writer;
  wave_force<ORDER, F64, P, NGB, PIPE>(rec, idr, ipos, ivel, iacc, ih2, iid, g * P, min(P, ni - g * P), jb, je, eps2d,
                                 ngb_cnt, ngb_list, T, dsd, nnaddr, islot, writer);
  if (blockred) {
    // The block's waves are splits 4q..4q+3 of one i-group (S % 4 == 0, so
    // every block is whole and no wave left early).  Their sums go through
    // LDS four fields at a time; wave w adds field c0 + w over the four
    // splits in the fixed order (s0 + s1) + (s2 + s3) and writes one row.
    // Wave 3 also takes the nearest neighbour.  The barriers bound what any
    // wave holds at once, so the epilogue does not raise the register count.
    __shared__ double red_t[4][4][64];
    __shared__ double red_d[4][64];
    __shared__ int red_a[4][64];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int i = g * P + islot;
    const int S4 = S >> 2, row = sidx >> 2;
    red_d[w][lane] = dsd;
    red_a[w][lane] = nnaddr;
#pragma unroll
    for (int c0 = 0; c0 < NF; c0 += 4) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c0 + k < NF) red_t[w][k][lane] = T[c0 + k];
      __syncthreads();
      if (c0 + w < NF && writer) {
        const double v = (red_t[0][w][lane] + red_t[1][w][lane]) + (red_t[2][w][lane] + red_t[3][w][lane]);
        part[((size_t)(c0 + w) * S4 + row) * ni + i] = v;
      }
      if (c0 == 0 && w == 3 && writer) {
        double bd = red_d[0][lane];
        int ba = red_a[0][lane];
#pragma unroll
        for (int u = 1; u < 4; ++u) {
          const double od = red_d[u][lane];
          const int oa = red_a[u][lane];
          const bool take = (od < bd) || (od == bd && oa < ba);
          bd = take ? od : bd;
          ba = take ? oa : ba;
        }
        part_ds2[(size_t)row * ni + i] = bd;
        part_addr[(size_t)row * ni + i] = ba;
      }
      __syncthreads();
    }
    return;
  }
  if (writer) {
    const int i = g * P + islot;
#pragma unroll
    for (int f = 0; f < NF; ++f) part[((size_t)f * S + sidx) * ni + i] = T[f];
    part_ds2[(size_t)sidx * ni + i] = dsd;
    part_addr[(size_t)sidx * ni + i] = nnaddr;
  }
}

template <int ORDER, bool F64>
__device__ __forceinline__ void write_row(const FinalizeArgs& a, int i, const double* T, double ds, int ad,
                                          int cnt);

// Ensemble force: one wave per item of a work table (cluster, i-group,
// j-split); each item carries its own lane shape.  Without lists, an item
// that is its i-particles' only split writes their result rows (over PCIe,
// while other items still compute) as the finalize would: the finalize's
// sum over one split is 0.0 + T, and its nearest neighbour the item's own.
template <int ORDER, bool F64, int P, bool NGB>
__device__ __forceinline__ void ens_item(const EnsArgs& a, const FinalizeArgs& fz,
                                         const typename RecOf<ORDER, F64>::T* __restrict__ rec, const EnsItem& it,
                                         int item) {
  constexpr int NF = nf<ORDER>();
  double T[NF];
  double dsd;
  int nnaddr, islot;
  bool writer;
  wave_force<ORDER, F64, P, NGB>(rec, a.idr, a.ipos, a.ivel, a.iacc, a.ih2, a.iid,
                                 it.i0, it.icount, it.jb, it.je, a.eps2[it.cluster], a.ngb_cnt, a.ngb_list, T,
                                 dsd, nnaddr, islot, writer);
  if (!NGB && it.nsplit == 1) {
    if (writer) {
#pragma unroll
      for (int f = 0; f < NF; ++f) T[f] = 0.0 + T[f];
      write_row<ORDER, F64>(fz, it.i0 + islot, T, dsd, nnaddr, 0);
    }
    return;
  }
  if (writer) {
    const size_t row = (size_t)item * 64 + islot;
#pragma unroll
    for (int f = 0; f < NF; ++f) a.part[(size_t)f * a.nitems * 64 + row] = T[f];
    a.part_ds2[row] = dsd;
    a.part_addr[row] = nnaddr;
  }
}

// P64: items [ib, ie) all have lane shape 64; otherwise 16, 4 or 1 (a switch
// whose larger register needs stay out of the P = 64 kernel).
// One single-wave work-group per position q of the ordered item list [ib, ie):
// the list holds the longest items first and the hardware dispatches work-
// groups in order as wave slots free up, so the short items fill the end of
// the launch with no queue to contend for.  (A counter that every wave bumped
// for its next item serialised the small items of large ensembles.)
template <int ORDER, bool F64, bool NGB, bool P64>
__global__ __launch_bounds__(64) void k_force_ens(EnsArgs a, FinalizeArgs fz,
                                                  const typename RecOf<ORDER, F64>::T* __restrict__ rec, int ib,
                                                  int ie) {
  const int q = ib + (int)blockIdx.x;
  if (q >= ie) return;
  const int item = __builtin_amdgcn_readfirstlane(a.order[q]);
  // the item is wave-uniform; readfirstlane keeps its fields (and so the
  // j-record addresses) in SGPRs, which a load the compiler cannot prove
  // invariant would otherwise place in VGPRs
  const EnsItem ld = a.items[item];
  EnsItem it;
  it.i0 = __builtin_amdgcn_readfirstlane(ld.i0);
  it.icount = __builtin_amdgcn_readfirstlane(ld.icount);
  it.jb = __builtin_amdgcn_readfirstlane(ld.jb);
  it.je = __builtin_amdgcn_readfirstlane(ld.je);
  it.P = __builtin_amdgcn_readfirstlane(ld.P);
  it.cluster = __builtin_amdgcn_readfirstlane(ld.cluster);
  it.nsplit = __builtin_amdgcn_readfirstlane(ld.nsplit);
  it.pad = 0;
  if constexpr (P64) {
    ens_item<ORDER, F64, 64, NGB>(a, fz, rec, it, item);
  } else {
    switch (it.P) {
      case 16: ens_item<ORDER, F64, 16, NGB>(a, fz, rec, it, item); break;
      case 4: ens_item<ORDER, F64, 4, NGB>(a, fz, rec, it, item); break;
      default: ens_item<ORDER, F64, 1, NGB>(a, fz, rec, it, item); break;
    }
  }
}

// Ensemble finalize, thread per i-particle: fixed-order sum over its items.
template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_finalize_ens(FinalizeArgs f, EnsArgs a) {
  constexpr int NF = nf<ORDER>();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= f.cd->ni) return;
  const int4 d = a.idesc[i];  // first item, number of splits, i-slot
  if (!f.ngb_on && d.y == 1) return;  // its only item wrote the row
  double T[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) T[q] = 0.0;
  double ds = INFINITY;
  int ad = INT_MAX;
  for (int s = 0; s < d.y; ++s) {
    const size_t row = (size_t)(d.x + s) * 64 + d.z;
#pragma unroll
    for (int q = 0; q < NF; ++q) T[q] += a.part[(size_t)q * a.nitems * 64 + row];
    const double dd = a.part_ds2[row];
    const int ad2 = a.part_addr[row];
    const bool take = (dd < ds) || (dd == ds && ad2 < ad);
    ds = take ? dd : ds;
    ad = take ? ad2 : ad;
  }
  write_row<ORDER, F64>(f, i, T, ds, ad, f.ngb_on ? f.ngb_cnt[i] : 0);
}

// ---------------------------------------------------------------------------
// Finalize.  Fixed-order sums of the S partial rows of each i-particle, the
// nearest neighbour as the (ds2, address) minimum, the result row, and the
// ascending sort of the neighbour list.  Two shapes:
//   k_finalize_rows  one thread per i-particle (partials read coalesced over i),
//                    followed by k_sort_lists (one wave per list);
//   k_finalize_block one work-group per i-particle, threads striding over the
//                    splits (few i-particles, many splits).
// Result row of i-particle i, field-major (consecutive i of a wave write
// consecutive doubles of each field: whole lines over PCIe).
template <int ORDER, bool F64>
__device__ __forceinline__ void write_row(const FinalizeArgs& a, int i, const double* T, double ds, int ad,
                                          int cnt) {
  using Rec = typename RecOf<ORDER, F64>::T;
  constexpr int NF = nf<ORDER>();
  double* o = a.out + i;
  const size_t st = (size_t)a.ostride;
#pragma unroll
  for (int f = 0; f < NF; ++f) o[f * st] = T[f];
  const bool has = ad != INT_MAX;
  // the device returns ds2 rounded to float (kernels.cu:999-1002)
  o[NF * st] = has ? (double)(float)ds : INFINITY;
  o[(NF + 1) * st] = has ? (double)reinterpret_cast<const Rec*>(a.rec)[ad].id : -1.0;
  o[(NF + 2) * st] = (double)cnt;
}

// Rank sort of c ids staged in s[0..c) by one wave, written to dst.
__device__ __forceinline__ void wave_rank_sort(const int* s, int c, int lane, int* dst) {
  for (int k = lane; k < c; k += 64) {
    const int v = s[k];
    int rank = 0;
    for (int q = 0; q < c; ++q) {
      const int u = s[q];
      rank += (u < v) || (u == v && q < k);
    }
    dst[rank] = v;
  }
}

// Four consecutive lanes per i-particle: lane u sums rows u, u + 4, ... in
// order, then the four sums combine as (S0 + S1) + (S2 + S3) (IEEE addition
// commutes exactly, so every lane of the four holds the same value).
template <int ORDER, bool F64>
__device__ __forceinline__ void finalize_rows_block(const FinalizeArgs& a, int b) {
  constexpr int NF = nf<ORDER>();
  const int ni = a.cd->ni;
  const int t = b * 256 + threadIdx.x;
  const int i = t >> 2, u = t & 3;
  if (i >= ni) return;  // the four lanes of an i-particle leave together
  double T[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) T[f] = 0.0;
  double ds = INFINITY;
  int ad = INT_MAX;
  for (int s = u; s < a.S; s += 4) {
#pragma unroll
    for (int f = 0; f < NF; ++f) T[f] += a.part[((size_t)f * a.S + s) * ni + i];
    const double d = a.part_ds2[(size_t)s * ni + i];
    const int ad2 = a.part_addr[(size_t)s * ni + i];
    const bool take = (d < ds) || (d == ds && ad2 < ad);
    ds = take ? d : ds;
    ad = take ? ad2 : ad;
  }
#pragma unroll
  for (int m = 1; m < 4; m <<= 1) {
#pragma unroll
    for (int f = 0; f < NF; ++f) T[f] += __shfl_xor(T[f], m);
    const double od = __shfl_xor(ds, m);
    const int oa = __shfl_xor(ad, m);
    const bool take = (od < ds) || (od == ds && oa < ad);
    ds = take ? od : ds;
    ad = take ? oa : ad;
  }
  if (u == 0) write_row<ORDER, F64>(a, i, T, ds, ad, a.ngb_on ? a.ngb_cnt[i] : 0);
}

template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_finalize_rows(FinalizeArgs a) {
  finalize_rows_block<ORDER, F64>(a, blockIdx.x);
}

// Host offset of i-particle i's list in a single-system call: its region's
// base plus the stored lengths of the region's earlier lists.  The counts are
// final once the force kernel has run, so no work-group waits for another; a
// whole wave calls this (at most LIST_REGION - 1 counts, a few per lane).
__device__ __forceinline__ int region_offset(const int* __restrict__ cnt, int i, int lane) {
  const int r0 = i & ~(LIST_REGION - 1);
  int s = 0;
  for (int k = r0 + lane; k < i; k += 64) s += min(cnt[k], NGB_CAP);
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m);
  return r0 * NGB_CAP + s;
}

// Writes the sorted list s[0..c) of i-particle i: back into its device slots
// and, when it fits, to the host-mapped buffer at off (consecutive lanes write
// consecutive ids).
__device__ __forceinline__ void store_list(const FinalizeArgs& a, int i, const int* s, int c, int off, int t,
                                           int nt) {
  const bool host = (long)off + c <= a.packed_cap;
  for (int k = t; k < c; k += nt) {
    const int v = s[k];
    a.ngb_list[(size_t)i * NGB_CAP + k] = v;
    if (host) a.packed[off + k] = v;
  }
}

// One wave per i-particle: ascending sort of its (at most NGB_CAP) list in LDS,
// then the host copy: in i order within its region, or (ensembles) packed, the
// four lists of a work-group taking their offsets with one atomic on the
// call's counter.
__device__ __forceinline__ void sort_lists_block(const FinalizeArgs& a, int b) {
  __shared__ int lst[4][NGB_CAP];
  __shared__ int srt[4][NGB_CAP];
  __shared__ int cnts[4];
  __shared__ int base;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int i = b * 4 + w;
  const bool valid = i < a.cd->ni;
  const int c = valid ? min(a.ngb_cnt[i], NGB_CAP) : 0;
  for (int k = lane; k < c; k += 64) lst[w][k] = a.ngb_list[(size_t)i * NGB_CAP + k];
  if (lane == 0) cnts[w] = c;
  __syncthreads();
  if (threadIdx.x == 0 && !a.inorder) {
    const int tot = cnts[0] + cnts[1] + cnts[2] + cnts[3];
    base = tot > 0 ? atomicAdd(a.pack_ctr, tot) : 0;
  }
  wave_rank_sort(lst[w], c, lane, srt[w]);
  int off = valid && a.inorder ? region_offset(a.ngb_cnt, i, lane) : 0;
  __syncthreads();
  if (!a.inorder) {
    off = base;
    for (int u = 0; u < w; ++u) off += cnts[u];
  }
  if (valid && lane == 0) a.list_off[i] = off;
  store_list(a, i, srt[w], c, off, lane, 64);
}

__global__ __launch_bounds__(256) void k_sort_lists(FinalizeArgs a) { sort_lists_block(a, blockIdx.x); }

// Finalize and list sort in one launch (independent work: the sort only
// reorders the lists, the finalize reads counts): the first nfin work-groups
// finalize, the others sort; both roles are uniform per work-group.
template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_finalize_rows_sort(FinalizeArgs a, int nfin) {
  if ((int)blockIdx.x < nfin) finalize_rows_block<ORDER, F64>(a, blockIdx.x);
  else sort_lists_block(a, blockIdx.x - nfin);
}

// One 256-thread block per i-particle: thread t sums the splits s = t, t+256, ...
// in order, then a fixed-order tree over the 256 threads in LDS (deterministic).
template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_finalize_block(FinalizeArgs a) {
  constexpr int NF = nf<ORDER>();
  __shared__ double shT[NF][256];
  __shared__ double shD[256];
  __shared__ int shA[256];
  __shared__ int lst[NGB_CAP];
  __shared__ int srt[NGB_CAP];
  __shared__ int base;
  const int t = threadIdx.x;
  const int i = blockIdx.x;
  const int ni = a.cd->ni;
  if (i >= ni) return;  // whole block (the grid may cover more i-particles)
  double T[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) T[f] = 0.0;
  double ds = INFINITY;
  int ad = INT_MAX;
  for (int s = t; s < a.S; s += 256) {
#pragma unroll
    for (int f = 0; f < NF; ++f) T[f] += a.part[((size_t)f * a.S + s) * ni + i];
    const double d = a.part_ds2[(size_t)s * ni + i];
    const int ad2 = a.part_addr[(size_t)s * ni + i];
    const bool take = (d < ds) || (d == ds && ad2 < ad);
    ds = take ? d : ds;
    ad = take ? ad2 : ad;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) shT[f][t] = T[f];
  shD[t] = ds;
  shA[t] = ad;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
#pragma unroll
      for (int f = 0; f < NF; ++f) shT[f][t] += shT[f][t + h];
      const double od = shD[t + h];
      const int oa = shA[t + h];
      if ((od < shD[t]) || (od == shD[t] && oa < shA[t])) { shD[t] = od; shA[t] = oa; }
    }
    __syncthreads();
  }
  const int cnt = a.ngb_on ? a.ngb_cnt[i] : 0;
  const int c = min(cnt, NGB_CAP);
  if (t == 0) {
    double R[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) R[f] = shT[f][0];
    write_row<ORDER, F64>(a, i, R, shD[0], shA[0], cnt);
    if (a.ngb_on && !a.inorder) {
      base = c > 0 ? atomicAdd(a.pack_ctr, c) : 0;
      a.list_off[i] = base;
    }
  }
  if (a.ngb_on && a.inorder && t < 64) {  // uniform over the block; a whole wave
    const int o = region_offset(a.ngb_cnt, i, t);
    if (t == 0) {
      base = o;
      a.list_off[i] = o;
    }
  }
  if (c == 0) return;  // uniform over the block
  for (int k = t; k < c; k += 256) lst[k] = a.ngb_list[(size_t)i * NGB_CAP + k];
  __syncthreads();
  if (t < 64) wave_rank_sort(lst, c, t, srt);
  __syncthreads();
  store_list(a, i, srt, c, base, t, 256);
}

// ---------------------------------------------------------------------------
// Predicted-state gather for read-back: per j, 10 doubles
// (x, y, z, m, vx, vy, vz, ax, ay, az) in FP64, plus the id in the m slot's
// companion array handled by the host.
template <int ORDER, bool F64>
__global__ __launch_bounds__(256) void k_gather_pred(GatherArgs a) {
  using Rec = typename RecOf<ORDER, F64>::T;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const int j = a.addr ? a.addr[k] : k;
  const Rec r = reinterpret_cast<const Rec*>(a.st.rec)[j];
  const size_t cap = (size_t)a.st.cap;
  double* o = a.out + (size_t)k * 10;
  o[3] = a.st.raw[RF_M * cap + j];
  if constexpr (F64) {
    o[0] = r.x; o[1] = r.y; o[2] = r.z;
    o[4] = r.vx; o[5] = r.vy; o[6] = r.vz;
    if constexpr (ORDER >= ORDER_H6) { o[7] = r.ax; o[8] = r.ay; o[9] = r.az; }
    else { o[7] = o[8] = o[9] = 0.0; }
  } else {
    o[0] = a.st.ppos[j]; o[1] = a.st.ppos[cap + j]; o[2] = a.st.ppos[2 * cap + j];
    o[4] = a.st.pvel[j]; o[5] = a.st.pvel[cap + j]; o[6] = a.st.pvel[2 * cap + j];
    if constexpr (ORDER >= ORDER_H6) { o[7] = a.st.pacc[j]; o[8] = a.st.pacc[cap + j]; o[9] = a.st.pacc[2 * cap + j]; }
    else { o[7] = o[8] = o[9] = 0.0; }
  }
}

__global__ void k_write_desc(CallDesc* cd, CallDesc v) { *cd = v; }

// ---------------------------------------------------------------------------
// Launchers.
static inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }
static inline int hip_status() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// Calls f(order constant, precision constant) for the runtime (order, f64).
template <typename F>
static void dispatch_op(int order, bool f64, F&& f) {
  using H4 = std::integral_constant<int, ORDER_H4>;
  using H6 = std::integral_constant<int, ORDER_H6>;
  if (order >= ORDER_H6) {
    if (f64) f(H6{}, std::true_type{}); else f(H6{}, std::false_type{});
  } else {
    if (f64) f(H4{}, std::true_type{}); else f(H4{}, std::false_type{});
  }
}

int launch_scatter(hipStream_t s, int order, bool f64, const ScatterArgs& a) {
  if (a.grid_blocks <= 0) return 0;
  const int nb = a.grid_blocks;  // grid-stride over the descriptor's counts
  dispatch_op(order, f64, [&](auto OR, auto FP) {
    hipLaunchKernelGGL((k_scatter<decltype(OR)::value, decltype(FP)::value>), dim3(nb), dim3(256), 0, s, a);
  });
  return hip_status();
}

int launch_predict(hipStream_t s, int order, bool f64, const PredictArgs& a) {
  if (a.j1 <= a.j0) return 0;
  const int nb = blocks_for(a.j1 - a.j0, 256);
  dispatch_op(order, f64, [&](auto OR, auto FP) {
    hipLaunchKernelGGL((k_predict<decltype(OR)::value, decltype(FP)::value>), dim3(nb), dim3(256), 0, s, a);
  });
  return hip_status();
}

bool force_p_supported(int P) {
  return P == 64 || P == 32 || P == 16 || P == 8 || P == 4 || P == 2 || P == 1;
}

template <int OR, bool FP, int P>
static void launch_force_p(hipStream_t s, const ForceArgs& a, const ForceConfig& c) {
  using Rec = typename RecOf<OR, FP>::T;
  const long G = (a.grid_ni + P - 1) / P;
  const int nb = blocks_for(G * c.S, 4);
  const Rec* rec = reinterpret_cast<const Rec*>(a.rec);
#define SAPK_FORCE_ARGS rec, a.idr, a.nj, a.ibuf, a.iid, a.cd, a.part, a.part_ds2, \
                        a.part_addr, a.ngb_cnt, a.ngb_list, c.S, c.jchunk, c.smajor, partial_rows(c) != c.S ? 1 : 0
  if (a.ngb_on && c.pipe)
    hipLaunchKernelGGL((k_force<OR, FP, P, true, true>), dim3(nb), dim3(256), 0, s, SAPK_FORCE_ARGS);
  else if (a.ngb_on)
    hipLaunchKernelGGL((k_force<OR, FP, P, true, false>), dim3(nb), dim3(256), 0, s, SAPK_FORCE_ARGS);
  else
    hipLaunchKernelGGL((k_force<OR, FP, P, false, false>), dim3(nb), dim3(256), 0, s, SAPK_FORCE_ARGS);
#undef SAPK_FORCE_ARGS
}

template <int OR, bool FP>
static int launch_force_t(hipStream_t s, const ForceArgs& a, const ForceConfig& c) {
  switch (c.P) {
    case 64: launch_force_p<OR, FP, 64>(s, a, c); break;
    case 32: launch_force_p<OR, FP, 32>(s, a, c); break;
    case 16: launch_force_p<OR, FP, 16>(s, a, c); break;
    case 8: launch_force_p<OR, FP, 8>(s, a, c); break;
    case 4: launch_force_p<OR, FP, 4>(s, a, c); break;
    case 2: launch_force_p<OR, FP, 2>(s, a, c); break;
    case 1: launch_force_p<OR, FP, 1>(s, a, c); break;
    default: return -2;
  }
  return hip_status();
}

int launch_force(hipStream_t s, int order, bool f64, const ForceArgs& a, const ForceConfig& c) {
  if (a.grid_ni <= 0 || a.nj <= 0) return 0;
  if (c.S <= 0 || c.jchunk <= 0 || (long)c.S * c.jchunk < a.nj) return -3;
  int rc = 0;
  dispatch_op(order, f64, [&](auto OR, auto FP) { rc = launch_force_t<decltype(OR)::value, decltype(FP)::value>(s, a, c); });
  return rc;
}

int launch_finalize(hipStream_t s, int order, bool f64, const FinalizeArgs& a) {
  if (a.grid_ni <= 0) return 0;
  // rows: coalesced over i, best when many i-particles share few splits
  dispatch_op(order, f64, [&](auto OR, auto FP) {
    constexpr int O = decltype(OR)::value;
    constexpr bool F = decltype(FP)::value;
    if (a.rows) {
      const int nfin = blocks_for(4L * a.grid_ni, 256);
      if (a.ngb_on)
        hipLaunchKernelGGL((k_finalize_rows_sort<O, F>), dim3(nfin + blocks_for(a.grid_ni, 4)), dim3(256), 0, s, a,
                           nfin);
      else
        hipLaunchKernelGGL((k_finalize_rows<O, F>), dim3(nfin), dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL((k_finalize_block<O, F>), dim3(a.grid_ni), dim3(256), 0, s, a);
    }
  });
  return hip_status();
}

int launch_force_ens(hipStream_t s, int order, bool f64, const EnsArgs& a, const FinalizeArgs& f, int n64) {
  if (a.nitems <= 0 || f.grid_ni <= 0) return 0;
  const int nsmall = a.nitems - n64;
  dispatch_op(order, f64, [&](auto OR, auto FP) {
    constexpr int O = decltype(OR)::value;
    constexpr bool F = decltype(FP)::value;
    const auto* rec = reinterpret_cast<const typename RecOf<O, F>::T*>(a.rec);
    if (n64 > 0) {
      if (a.ngb_on)
        hipLaunchKernelGGL((k_force_ens<O, F, true, true>), dim3(n64), dim3(64), 0, s, a, f, rec, 0, n64);
      else
        hipLaunchKernelGGL((k_force_ens<O, F, false, true>), dim3(n64), dim3(64), 0, s, a, f, rec, 0, n64);
    }
    if (nsmall > 0) {
      if (a.ngb_on)
        hipLaunchKernelGGL((k_force_ens<O, F, true, false>), dim3(nsmall), dim3(64), 0, s, a, f, rec, n64, a.nitems);
      else
        hipLaunchKernelGGL((k_force_ens<O, F, false, false>), dim3(nsmall), dim3(64), 0, s, a, f, rec, n64, a.nitems);
    }
    hipLaunchKernelGGL((k_finalize_ens<O, F>), dim3(blocks_for(f.grid_ni, 256)), dim3(256), 0, s, f, a);
    if (f.ngb_on) hipLaunchKernelGGL(k_sort_lists, dim3(blocks_for(f.grid_ni, 4)), dim3(256), 0, s, f);
  });
  return hip_status();
}

// One thread per i-particle: its group by binary search over the groups' first
// i-particles (groups are in i order and cover [0, ni)).
__global__ __launch_bounds__(256) void k_ens_idesc(const EnsGroup* __restrict__ g, int ng, int ni,
                                                   int4* __restrict__ idesc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ni) return;
  int lo = 0, hi = ng - 1;  // the last group with ib <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].ib <= i) lo = mid;
    else hi = mid - 1;
  }
  const EnsGroup gr = g[lo];
  const int r = i - gr.ib;
  idesc[i] = make_int4(gr.first + (r / gr.P) * gr.S, gr.S, r % gr.P, gr.P == 64 ? 0 : 1);
}

int launch_ens_idesc(hipStream_t s, const EnsGroup* g, int ng, int ni, int4* idesc) {
  if (ni <= 0 || ng <= 0) return 0;
  hipLaunchKernelGGL(k_ens_idesc, dim3(blocks_for(ni, 256)), dim3(256), 0, s, g, ng, ni, idesc);
  return hip_status();
}

int launch_gather_pred(hipStream_t s, int order, bool f64, const GatherArgs& a) {
  if (a.n <= 0) return 0;
  const int nb = blocks_for(a.n, 256);
  dispatch_op(order, f64, [&](auto OR, auto FP) {
    hipLaunchKernelGGL((k_gather_pred<decltype(OR)::value, decltype(FP)::value>), dim3(nb), dim3(256), 0, s, a);
  });
  return hip_status();
}

int launch_write_desc(hipStream_t s, CallDesc* cd, const CallDesc& v) {
  hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(1), 0, s, cd, v);
  return hip_status();
}

}  // namespace sapk
