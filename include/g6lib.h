/*
 * g6lib.h -- GRAPE-6 compatible C-ABI of libsapporo_amd.so (Hermite-4).
 *
 * Drop-in replacement for the reference's GRAPE-6 emulation library
 * (lib/interfaces/sapporoG6lib.cpp:11-205) and its caller header
 * (testCodes/src/g6lib.h:4-90).  Every symbol keeps the reference's name,
 * argument order, units and return values; a Hermite-4 integrator linked
 * against the reference links against this library unchanged.
 *
 * Units and conventions (sapporoG6lib.cpp:89-108, kernels.cu:565-669):
 *   G = 1; g6_set_j_particle_ takes a/2 (a2) and j/6 (j6), scaled x2 and x6
 *   inside; pot is negative (-sum m/r); inn is the nearest neighbour's id;
 *   neighbour lists are ids sorted ascending; eps2 is global per call.
 * Precision: the reference opens FOURTH/DOUBLESINGLE (sapporoG6lib.cpp:34).
 *   Here the default is the mixed-precision kernel (positions as double-single
 *   pairs, relative positions by the double-single difference
 *   (xj.hi - xi.hi) + (xj.lo - xi.lo) in FP32, FP32 pair terms, per-tile FP32
 *   sums folded into FP64); SAPPORO2_AMD_PRECISION=double selects the all-FP64
 *   kernel.  Both ignore cluster_id, as the reference.
 *
 * The second part of this header is the caller side of the reference header
 * (testCodes/src/g6lib.h:52-90): by-value wrappers, the position-only
 * multisend helper and the no-op stubs.  The reference defines them as plain
 * (external) functions in its header; here they are static inline, so the
 * header can be included by several translation units of one program.
 */
#ifndef SAPPORO2_AMD_G6LIB_H
#define SAPPORO2_AMD_G6LIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* sapporoG6lib.cpp:25  open with an explicit device list (first entry used). */
int g6_open_special(int ndev, int *list);
/* sapporoG6lib.cpp:46  reads ./sapporo2.config (ndev, then device ids). */
int g6_open_(int *id);
/* sapporoG6lib.cpp:80 */
int g6_close_(int *id);
/* sapporoG6lib.cpp:81  returns the number of i-particles per call: 16384. */
int g6_npipes_(void);
/* sapporoG6lib.cpp:86-87  no-ops, return 0. */
int g6_set_tunit_(double *);
int g6_set_xunit_(double *);
/* sapporoG6lib.cpp:88  prediction time of the next force call. */
int g6_set_ti_(int *id, double *ti);
/* sapporoG6lib.cpp:89-108  stores particle `address` (addresses must arrive
 * in order: capacity grows by 10%); k18 is ignored. */
int g6_set_j_particle_(int *cluster_id, int *address, int *index, double *tj, double *dtj,
                       double *mass, double k18[3], double j6[3], double a2[3], double v[3],
                       double x[3]);
/* sapporoG6lib.cpp:109-120  starts the force on ni i-particles from the nj
 * j-addresses [0, nj); returns without waiting.  aold/j6old/phiold unused. */
void g6calc_firsthalf_(int *cluster_id, int *nj, int *ni, int index[], double xi[][3],
                       double vi[][3], double aold[][3], double j6old[][3], double phiold[3],
                       double *eps2, double h2[]);
/* sapporoG6lib.cpp:121-134  waits and writes acc, jerk, pot. */
int g6calc_lasthalf_(int *cluster_id, int *nj, int *ni, int index[], double xi[][3],
                     double vi[][3], double *eps2, double h2[], double acc[][3],
                     double jerk[][3], double pot[]);
/* sapporoG6lib.cpp:135-149  as lasthalf plus nearest-neighbour ids in inn
 * (0 when none closer than ds2 = 1e10). */
int g6calc_lasthalf2_(int *cluster_id, int *nj, int *ni, int index[], double xi[][3],
                      double vi[][3], double *eps2, double h2[], double acc[][3],
                      double jerk[][3], double pot[], int *inn);
/* sapporoG6lib.cpp:176-179  no-ops, return 0. */
int g6_initialize_jp_buffer_(int *cluster_id, int *buf_size);
int g6_flush_jp_buffer_(int *cluster_id);
int g6_reset_(int *cluster_id);
int g6_reset_fofpga_(int *cluster_id);
/* sapporoG6lib.cpp:182  fetches the lists of the last call; returns 1 if any
 * i-particle has >= 256 neighbours (contents then unspecified). */
int g6_read_neighbour_list_(int *cluster_id);
/* sapporoG6lib.cpp:187-201  list of i-particle ipipe: up to maxlength ids,
 * sorted ascending, n_neighbours = count; returns 1 if count >= maxlength. */
int g6_get_neighbour_list_(int *cluster_id, int *ipipe, int *maxlength, int *n_neighbours,
                           int neighbour_list[]);

/* sapporoG6lib.cpp:152-172  raw and predicted state of j-particle addr. */
#ifdef __cplusplus
void get_j_part_data(int addr, int nj, double *pos, double *vel, double *acc, double *jrk,
                     double *ppos, double *pvel, double &mass, double &eps2, int &id);
#else
void get_j_part_data(int addr, int nj, double *pos, double *vel, double *acc, double *jrk,
                     double *ppos, double *pvel, double *mass, double *eps2, int *id);
#endif

/* ---- caller side: testCodes/src/g6lib.h:52-62, by-value wrappers ---- */
static inline int g6_close(int clusterid) { return g6_close_(&clusterid); }
static inline int g6_open(int clusterid) { return g6_open_(&clusterid); }
static inline int g6_npipes(void) { return g6_npipes_(); }
static inline void g6_set_ti(int cluster_id, double ti) { g6_set_ti_(&cluster_id, &ti); }
static inline int g6_reset(int cluster_id) { return g6_reset_(&cluster_id); }
static inline int g6_reset_fofpga(int cluster_id) { return g6_reset_fofpga_(&cluster_id); }
static inline int g6_set_j_particle(int clusterid, int address, int index, double tj, double dtj, double mass,
                                    double k18[3], double j6[3], double a2[3], double v[3], double x[3]) {
  return g6_set_j_particle_(&clusterid, &address, &index, &tj, &dtj, &mass, k18, j6, a2, v, x);
}
static inline void g6calc_firsthalf(int cluster_id, int nj, int ni, int index[], double xi[][3], double vi[][3],
                                    double aold[][3], double j6old[][3], double phiold[], double eps2,
                                    double h2[]) {
  g6calc_firsthalf_(&cluster_id, &nj, &ni, index, xi, vi, aold, j6old, phiold, &eps2, h2);
}
static inline int g6calc_lasthalf(int cluster_id, int nj, int ni, int index[], double xi[][3], double vi[][3],
                                  double eps2, double h2[], double acc[][3], double jerk[][3], double pot[]) {
  return g6calc_lasthalf_(&cluster_id, &nj, &ni, index, xi, vi, &eps2, h2, acc, jerk, pot);
}
static inline void g6_set_tunit(double newtunit) { (void)g6_set_tunit_(&newtunit); }
static inline void g6_set_xunit(double newxunit) { (void)g6_set_xunit_(&newxunit); }

/* testCodes/src/g6lib.h:67-84: stores particle `address` with its mass and
 * position only (time, velocity, acceleration and jerk zero). */
static inline int g6_set_j_particle_multisend_mxfast_(int *clusterid, int *nclusters, int *address, int *index,
                                                      double *mass, double x[3]) {
  double zero3a[3] = {0, 0, 0}, zero3b[3] = {0, 0, 0}, zero3c[3] = {0, 0, 0}, zero3d[3] = {0, 0, 0};
  double tj = 0, dtj = 0;
  (void)nclusters;
  return g6_set_j_particle_(clusterid, address, index, &tj, &dtj, mass, zero3a, zero3b, zero3c, zero3d, x);
}

/* testCodes/src/g6lib.h:86-90: hardware-management calls of GRAPE-6 boards,
 * no-ops here (the reference's bodies return nothing; these return 0). */
static inline int g6_reinitialize(int clusterid) { (void)clusterid; return 0; }
static inline int g6_print_chip_status(int clusterid) { (void)clusterid; return 0; }
static inline int g6_flush_jp_buffer_and_multisend(int clusterid, int something) {
  (void)clusterid; (void)something; return 0;
}
static inline int g6_initialize_jp_buffer(int clusterid, int list_max) { (void)clusterid; (void)list_max; return 0; }
static inline int g6_set_overflow_flag_test_mode_(int *clusterid) { (void)clusterid; return 0; }

#ifdef __cplusplus
}
#endif

#endif /* SAPPORO2_AMD_G6LIB_H */
