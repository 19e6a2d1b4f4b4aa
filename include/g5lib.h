/* g5lib.h -- GRAPE-5 C-ABI of sapporo2_amd (libsapporo_amd.so).
 *
 * The functions of the reference's lib/interfaces/sapporoG5lib.cpp:40-218 with
 * the same names, arguments and behaviour: second-order force (acceleration and
 * potential) in single precision, no self-exclusion (a particle's own mass adds
 * -m/eps to its potential, as on GRAPE-5 hardware), no predictor, no
 * neighbour information.  One global system per process, as in the reference.
 */
#ifndef SAPPORO2_AMD_G5LIB_H
#define SAPPORO2_AMD_G5LIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* sapporoG5lib.cpp:40-47: open on the listed devices (the first one is used) */
int g5_open_special(int ndev, int *list);
/* sapporoG5lib.cpp:49-82: open on the device named by ./sapporo2.config */
void g5_open(void);
/* sapporoG5lib.cpp:84-87 */
void g5_close(void);
/* sapporoG5lib.cpp:89-92: i-particles per call (16384) */
int g5_get_number_of_pipelines(void);
/* sapporoG5lib.cpp:94-97: number of j-particles [0, nj) used */
void g5_set_n(int nj);
/* sapporoG5lib.cpp:99-102: softening eps^2 for every pair */
void g5_set_eps2_to_all(double eps2);
/* sapporoG5lib.cpp:104-107: softening eps for every pair (squared here) */
void g5_set_eps_to_all(double eps);
/* sapporoG5lib.cpp:109-119: j-particles [addr, addr + nj) from xj[addr..], jmass[addr..] */
void g5_set_xmj(int addr, int nj, double xj[][3], double jmass[]);
/* sapporoG5lib.cpp:121-125 */
void g5_set_jp(int addr, int nj, double jmass[], double xj[][3]);
/* sapporoG5lib.cpp:127-150: start the force on ni <= 16384 i-particles */
void g5_set_xi(int ni, double xi[][3]);
/* sapporoG5lib.cpp:152: no-op (the calculation starts in g5_set_xi) */
void g5_run(void);
/* sapporoG5lib.cpp:154-161: wait for and return the force started by g5_set_xi */
void g5_get_force(int ni, double a[][3], double *phi);
/* sapporoG5lib.cpp:163-195: force on ni i-particles (any ni, in chunks of 16384) */
void g5_calculate_force_on_x(double xi[][3], double a[][3], double *phi, int ni);
/* sapporoG5lib.cpp:197-202: no-op on a GPU */
void g5_set_range(double xmin, double xmax, double mmin);
/* sapporoG5lib.cpp:204-218 */
int g5_get_jmemsize(void);
int g5_get_number_of_real_pipelines(void);
int g5_get_pcibus_freq(void);

#ifdef __cplusplus
}
#endif

#endif
