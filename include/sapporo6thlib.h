/*
 * sapporo6thlib.h -- 6th-order Hermite C-ABI of libsapporo6th_amd.so.
 *
 * Drop-in replacement for the reference's 6th-order interface
 * (lib/interfaces/sapporo6thlib.cpp:5-203): same symbols, argument order
 * and semantics.  acc/jrk/snp/crk are true (unscaled) derivatives; forces
 * return acc, jerk, snap and pot; crackle is returned as zero; neighbour
 * lists are not computed (h2 = 0), the nearest neighbour is.
 * Precision: the reference opens SIXTH/DOUBLE (sapporo6thlib.cpp:29); that is
 * the default here, SAPPORO2_AMD_PRECISION=mixed selects the mixed kernel.
 */
#ifndef SAPPORO2_AMD_6THLIB_H
#define SAPPORO2_AMD_6THLIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* sapporo6thlib.cpp:18-40  open once with a device list (first entry used). */
void initialize_special(int ndev, int *list);
/* sapporo6thlib.cpp:42-70  reads ./sapporo2.config if present. */
void initialize(void);
/* sapporo6thlib.cpp:84-93  store particle at address add. */
void set_j_particle(int add, double pos[3], double vel[3], double acc[3], double jrk[3],
                    double snp[3], double crk[3], double mass, double time, int id, double eps2);
/* sapporo6thlib.cpp:103-111  predict addresses [0, nj) to time. */
void predict_all(double time, int nj);
/* sapporo6thlib.cpp:119-127  flush updates without predicting. */
void no_predict_all(double time, int nj);
/* sapporo6thlib.cpp:143-149  predicted state of the particle at addr. */
#ifdef __cplusplus
void pick_up_predictor_2(int addr, int &id, double &mass, double &eps2, double pos[3],
                         double vel[3], double acc[3]);
#else
void pick_up_predictor_2(int addr, int *id, double *mass, double *eps2, double pos[3],
                         double vel[3], double acc[3]);
#endif
/* sapporo6thlib.cpp:167-201  forces on ni i-particles (any ni; chunked by
 * n_pipes) from the nj j-addresses [0, nj); global eps2 = eps2[0]. */
void calc_force_on_predictors(int ni, int nj, int ids[], double pos[][3], double vel[][3],
                              double acc[][3], double mass[], double eps2[], double accNew[][3],
                              double jrkNew[][3], double snpNew[][3], double crkNew[][3],
                              double phi[], int nnb[], double nnb_r2[]);

#ifdef __cplusplus
}
#endif

#endif /* SAPPORO2_AMD_6THLIB_H */
