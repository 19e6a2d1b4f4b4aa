/*
 * sapporo_amd.h -- engine-level C-ABI of libsapporo_amd.so.
 *
 * Handle-based form of the reference's host class `sapporo`
 * (lib/include/sapporohostclass.h:121-176): one handle per engine, so one
 * process can hold several systems (e.g. the mixed and the all-FP64 kernel
 * side by side).  The g6 and 6th-order ABIs are thin layers over it.
 * Arrays are row-major doubles (n x 3 for vectors); all pointers are plain
 * host pointers owned by the caller.  Returns: 0 on success, -1 on error
 * (message in sapamd_last_error) unless stated.
 *
 * order: 1 = FOURTH (Hermite-4), 2 = SIXTH (Hermite-6)  (defines.h:30)
 * precision: 1 = DOUBLESINGLE (mixed), 2 = DOUBLE (all FP64) (defines.h:31)
 */
#ifndef SAPPORO2_AMD_H
#define SAPPORO2_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sapamd_engine sapamd_engine;

/* New / delete an engine handle (no reference counterpart: the reference
 * holds one global instance, sapporoG6lib.cpp:3). */
sapamd_engine *sapamd_create(void);
void sapamd_destroy(sapamd_engine *e);

/* sapporohostclass.h:125  open(kernelFile, devices, nprocs, order, precision). */
int sapamd_open(sapamd_engine *e, int device, int order, int precision);
/* sapporohostclass.h:127 */
int sapamd_close(sapamd_engine *e);
/* sapporohostclass.h:128  i-particles per call (16384). */
int sapamd_get_n_pipes(sapamd_engine *e);
/* sapporohostclass.h:129-130 */
int sapamd_set_time(sapamd_engine *e, double t);
int sapamd_set_no_time(sapamd_engine *e);
/* sapporohostclass.h:132-139  set_j_particle(address, index, tj, dtj, mass,
 * k18, j6, a2, v, x, snp, crk, eps): acc and jrk are the true derivatives
 * (the g6 layer scales a/2 and j/6 before calling). snp/crk may be NULL. */
int sapamd_set_j_particle(sapamd_engine *e, int address, int id, double tj, double dtj,
                          double mass, const double acc[3], const double jrk[3],
                          const double vel[3], const double pos[3], const double snp[3],
                          const double crk[3], double eps);
/* Array form of sapamd_set_j_particle for n particles (no reference
 * counterpart; the reference takes one particle per call).  Vectors are
 * n x 3; tj, dtj, snp, crk and eps may be NULL (zeros). */
int sapamd_set_j_particles(sapamd_engine *e, int n, const int *address, const int *id,
                           const double *tj, const double *dtj, const double *mass,
                           const double *acc, const double *jrk, const double *vel,
                           const double *pos, const double *snp, const double *crk,
                           const double *eps);
/* sapporohostclass.h:141-147  startGravCalc; xi, vi, ai are ni x 3; ai may be
 * NULL for FOURTH; h2 may be NULL (no neighbour lists). */
int sapamd_start_grav_calc(sapamd_engine *e, int nj, int ni, const int *id, const double *xi,
                           const double *vi, const double *ai, double eps2, const double *h2);
/* sapporohostclass.h:149-156  getGravResults; any output may be NULL;
 * ngb != 0 requests nnb / dsmin (written only for a neighbour below 1e10). */
int sapamd_get_grav_results(sapamd_engine *e, int nj, int ni, double *acc, double *jrk,
                            double *snp, double *crk, double *pot, int *nnb, double *dsmin,
                            int ngb);
/* Ensembles of independent clusters on one engine (no reference counterpart:
 * the reference ignores cluster_id, sapporoG6lib.cpp:99, and holds one system
 * per process).  set_clusters lays out n clusters of nj[c] j-particles;
 * cluster c owns the addresses [cluster_offset(c), cluster_offset(c) + nj[c])
 * of set_j_particle.  start_grav_calc_clusters evaluates ni i-particles
 * grouped by cluster (cluster[] non-decreasing), each against its own
 * cluster only with softening eps2[cluster]; results, nearest neighbours and
 * lists are read with sapamd_get_grav_results / sapamd_get_ngb_list.
 * ni <= 4194304 per call. */
int sapamd_set_clusters(sapamd_engine *e, int n, const int *nj);
int sapamd_cluster_offset(sapamd_engine *e, int c);
int sapamd_start_grav_calc_clusters(sapamd_engine *e, int ni, const int *cluster, const int *id,
                                    const double *xi, const double *vi, const double *ai,
                                    const double *eps2, const double *h2);
/* sapporohostclass.h:170-171  returns 1 on overflow (a count >= 256). */
int sapamd_read_ngb_list(sapamd_engine *e);
/* sapporohostclass.h:172-176  returns 1 when the count >= maxlen. */
int sapamd_get_ngb_list(sapamd_engine *e, int ipipe, int maxlen, int *n, int *list);
/* sapporohostclass.h:158 */
int sapamd_force_prediction(sapamd_engine *e, int nj);
/* sapporohostclass.h:159-162 */
int sapamd_retrieve_predicted_j_particle(sapamd_engine *e, int addr, double *mass, double *id,
                                         double *eps2, double *pos, double *vel, double *acc);
/* sapporohostclass.h:164-168 */
int sapamd_retrieve_j_particle_state(sapamd_engine *e, int addr, double *mass, double *id,
                                     double *eps2, double *pos, double *vel, double *acc,
                                     double *jrk, double *ppos, double *pvel, double *pacc);

/* Launch shape and instrumentation (no reference counterpart; the reference
 * fixes its shape in sapporohostclass.cpp:1499-1561). */
int sapamd_set_force_override(sapamd_engine *e, int P, int S);
/* on: the 4 waves of a work-group share one j-range; off (default, faster on
 * MI355X): they take consecutive j-splits of one i-group */
int sapamd_set_split_major(sapamd_engine *e, int on);
/* on: neighbour-id stores are deferred by one group of pairs so the slot
 * atomic's latency overlaps arithmetic (more registers); off: immediate */
int sapamd_set_ngb_pipeline(sapamd_engine *e, int on);
int sapamd_set_timing(sapamd_engine *e, int on);
/* on: each force call's device work (scatter, predictor, force, finalize) is
 * captured as a hipGraph once per launch shape and replayed; the per-call
 * values travel in a device-side descriptor.  Results are identical to the
 * direct launches; per-kernel timings (last_force_ms, last_predict_ms) are 0.
 * No reference counterpart (the reference launches each kernel, sapporohostclass.cpp:1442-1602).
 * Default: off, or $SAPPORO2_AMD_GRAPHS. */
int sapamd_set_graphs(sapamd_engine *e, int on);
/* Checkpoint of the j-store, no counterpart in the reference (whose state
 * lives only in the process): sapamd_state_bytes flushes staged updates and
 * returns the size of a dump (-1 on error); sapamd_save_state writes it into a
 * caller-owned host buffer of at least that size; sapamd_load_state replaces
 * the whole store of a handle opened with the same order and precision (stored
 * fields, predicted state, block time, pending prediction, cluster layout), so
 * that the next calls give bit-identical results.  Results and neighbour lists
 * of earlier calls are not part of the dump. */
long long sapamd_state_bytes(sapamd_engine *e);
int sapamd_save_state(sapamd_engine *e, void *buf, long long bytes);
int sapamd_load_state(sapamd_engine *e, const void *buf, long long bytes);
/* Array form of sapamd_retrieve_predicted_j_particle for n addresses: mass,
 * predicted position, velocity (and acceleration, SIXTH; zero otherwise) of
 * each, gathered on the device so that only those rows cross PCIe.  No
 * counterpart in the reference, which copies every j back
 * (sapporohostclass.cpp:980-995).  Any output may be NULL. */
int sapamd_retrieve_predicted_j_particles(sapamd_engine *e, int n, const int *addr, double *mass,
                                          double (*pos)[3], double (*vel)[3], double (*acc)[3]);
int sapamd_graph_count(sapamd_engine *e);
float sapamd_last_force_ms(sapamd_engine *e);
float sapamd_last_predict_ms(sapamd_engine *e);
/* j-update scatter kernel time of the last call (timing on; 0 when the call had
 * no updates) and, in *nupd (may be NULL), the number of updates it applied */
float sapamd_last_scatter_ms(sapamd_engine *e, int *nupd);
/* finalize kernel time of the last call (timing on): reduction of the partial
 * sums, list sort, and the result and list writes into host memory */
float sapamd_last_finalize_ms(sapamd_engine *e);
/* device time of the last force call from the first kernel after its input
 * copies to the end of the finalize kernel (copies excluded) */
float sapamd_last_device_ms(sapamd_engine *e);
int sapamd_last_config(sapamd_engine *e, int *P, int *S, int *jchunk);
/* Launch-shape table (sapporo2_amd/autotune.py output; loaded at open from
 * $SAPPORO2_AMD_TUNE or tune/<arch>.txt beside the library).  Returns the
 * number of entries read (0: heuristic shapes). */
int sapamd_load_tuning(sapamd_engine *e, const char *path);
int sapamd_tuning_entries(sapamd_engine *e);
int sapamd_num_cus(sapamd_engine *e);
int sapamd_synchronize(sapamd_engine *e);
/* The engine behind the GRAPE-6 ABI's global instance (include/g6lib.h), so
 * that a g6 caller can switch on timing or read launch shapes with the calls
 * above.  No reference counterpart (the reference's instance is file-static,
 * sapporoG6lib.cpp:3).  Valid for the life of the process; do not destroy. */
sapamd_engine *sapamd_g6_engine(void);

/* Allocation-free variant of sapamd_start_grav_calc + sapamd_get_grav_results,
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  No counterpart in the reference.  It calls no hipMalloc,
 * hipFree or allocating routine and never waits on the host: the work is
 * queued on `stream` (a hipStream_t; NULL = the engine's own stream) and
 * complete when that stream is.
 *   id   : device pointer, ni ids
 *   ibuf : device pointer, [xi | vi | ai (SIXTH only)], each ni x 3 doubles
 *   out  : device pointer, field-major ni-strided rows: acc x, y, z, pot,
 *          jerk x, y, z, [snap x, y, z,] nearest-neighbour ds2 (float-rounded;
 *          +inf when none), nearest-neighbour id (-1 when none), 0
 * No neighbour lists.  0 < ni <= 16384 and nj within the j-store; staged
 * j-updates must have been flushed (sapamd_force_prediction) before, and a
 * pending prediction (sapamd_set_time) runs inside the call.  The call
 * descriptor and partial sums use a per-GPU device array of
 * sapamd_noalloc_scratch_bytes() bytes that the library holds until the
 * process ends; a call whose partial sums would not fit coarsens its j-split
 * (results then differ from sapamd_start_grav_calc by rounding only), and
 * one that still does not fit fails with -1.  All calls on one GPU share the
 * array: calls that overlap on one GPU race.  Returns 0 or -1. */
int sapamd_grav_calc_noalloc(sapamd_engine *e, int nj, int ni, const int *id, const double *ibuf,
                             double eps2, double *out, void *stream);
long long sapamd_noalloc_scratch_bytes(void);
const char *sapamd_last_error(sapamd_engine *e);

#ifdef __cplusplus
}
#endif

#endif /* SAPPORO2_AMD_H */
