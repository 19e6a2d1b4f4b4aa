/*
 * oracle.c -- CPU restatement of the reference's direct-summation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The results of the GPU path are compared with this
 * file's: only tests/, the smoke() entry and bench.py's cpu_baseline leg may load it.  The product
 * library (sapporo2_amd/) never links or calls it.
 *
 * Contents (each restates the reference file:line named beside it):
 *   orc_predict         FP64 Horner predictor        lib/CUDAKernels/kernels.cu:332-383,
 *                                                    lib/src/sapporohostclass.cpp:1304-1348
 *   orc_force           FP64 acc/pot/jerk[/snap] +   lib/src/sapporohostclass.cpp:1373-1438,
 *                       nearest neighbour + lists    lib/CUDAKernels/kernels.cu:565-667,
 *                       (FP64 working precision)     lib/interfaces/sapporoYeblib.cpp:57-86
 *   orc_neighbours_mixed nearest neighbour + lists   lib/CUDAKernels/kernels.cu:539-546, 565-589
 *                       with the mixed-precision     (DS kernel: FP32 ds2 and FP32 h2)
 *                       (FP32) distance arithmetic
 *
 * Arithmetic contract (shared with the HIP kernels, which must match it bit
 * for bit for the predictor, the neighbour lists and the nearest neighbour):
 *   predictor : the reference expression evaluated in IEEE double with no
 *               contraction (build with -ffp-contract=off).
 *   FP64 ds2  : dx = xj - xi (double); ds2 = fma(dz,dz, fma(dy,dy, dx*dx)).
 *   mixed ds2 : positions as double-single pairs, hi = (float)x,
 *               lo = (float)(x - hi) (kernels.cu:104-109, 163-167);
 *               dx = (xj.hi - xi.hi) + (xj.lo - xi.lo) in float (kernels.cu:170-178);
 *               ds2 = fmaf(dz,dz, fmaf(dy,dy, dx*dx)) (kernels.cu:567-570 with
 *               multiply-add contraction); h2 rounded to float.
 *   pair kept : id_i != id_j                         (kernels.cu:565)
 *   neighbour : h2_i > ds2 (strict)                  (kernels.cu:584)
 *   NN        : smallest ds2 (strict <) over kept pairs; exact ties go to the
 *               smallest j address (the reference's first-in-iteration rule,
 *               kernels.cu:541 and 868; equal to smallest id when ids increase
 *               with address).
 *   lists     : ids sorted ascending (sapporohostclass.cpp:871); the count is
 *               exact; at most `cap` ids are stored (the smallest ones).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORDER_G5 0
#define ORDER_H4 1
#define ORDER_H6 2

/* Predictor: reference kernels.cu:332-383 (same expression in
 * sapporohostclass.cpp:1304-1348).  For order <= FOURTH the reference sets
 * dt4 = dt5 = 0 and snap = crackle = 0 (kernels.cu:338-356); the zero terms
 * are kept so the IEEE operation sequence is the reference's.  Arrays are
 * n x 3 row-major; s, c, pa may be NULL for order < SIXTH. */
void orc_predict(int n, int order, double ti, const double *tj,
                 const double *x, const double *v, const double *a,
                 const double *jk, const double *s, const double *c,
                 double *px, double *pv, double *pa)
{
    for (int k = 0; k < n; ++k) {
        double dt = ti - tj[k];
        double dt2 = (1. / 2.) * dt;
        double dt3 = (1. / 3.) * dt;
        double dt4 = (1. / 4.) * dt;
        double dt5 = (1. / 5.) * dt;
        if (order <= ORDER_H4) {
            dt4 = 0;
            dt5 = 0;
        }
        for (int d = 0; d < 3; ++d) {
            const double xs = x[3 * k + d], vs = v[3 * k + d];
            const double as = a[3 * k + d], js = jk[3 * k + d];
            const double ss = (order >= ORDER_H6) ? s[3 * k + d] : 0.0;
            const double cs = (order >= ORDER_H6) ? c[3 * k + d] : 0.0;
            px[3 * k + d] = xs + dt * (vs + dt2 * (as + dt3 * (js + dt4 * (ss + dt5 * (cs)))));
            pv[3 * k + d] = vs + dt * (as + dt2 * (js + dt3 * (ss + dt4 * (cs))));
            if (order >= ORDER_H6)
                pa[3 * k + d] = as + dt * (js + dt2 * (ss + dt3 * (cs)));
        }
    }
}

static int cmp_int(const void *p, const void *q)
{
    const int a = *(const int *)p, b = *(const int *)q;
    return (a > b) - (a < b);
}

/* Keep the `cap` smallest ids seen so far in buf[0..*stored) (unsorted
 * insertion, replacing the current maximum when full). */
static void keep_smallest(int *buf, int *stored, int cap, int id)
{
    if (cap <= 0) return;
    if (*stored < cap) {
        buf[(*stored)++] = id;
        return;
    }
    int imax = 0;
    for (int k = 1; k < cap; ++k)
        if (buf[k] > buf[imax]) imax = k;
    if (id < buf[imax]) buf[imax] = id;
}

/* Full FP64 evaluation on ni i-particles against nj (predicted) j-particles.
 * Forces restate sapporohostclass.cpp:1388-1427 (acc, pot, jerk; inv =
 * 1/sqrt(ds2+eps2), self-pair skipped by id at 1394) and, for SIXTH, the
 * snap of kernels.cu:642-667 / sapporoYeblib.cpp:57-86.  Outputs (row-major
 * ni x 3 for vectors): acc, jrk, snp (SIXTH only, may be NULL), pot.
 * Nearest neighbour and neighbour lists use the FP64 distance contract in
 * the header; nn_addr = -1 when no kept pair exists.  ngb_list holds
 * ni x cap ids (sorted, smallest `cap` ids); ngb_cnt the exact count.
 * h2i may be NULL (no lists). */
void orc_force(int order, int ni, const double *xi, const double *vi,
               const double *ai, const double *h2i, const int *idi,
               int nj, const double *xj, const double *vj, const double *aj,
               const double *mj, const int *idj, double eps2,
               double *acc, double *jrk, double *snp, double *pot,
               int *nn_addr, double *nn_ds2, int *ngb_cnt, int *ngb_list,
               int cap, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; ++i) {
        const double px = xi[3 * i], py = xi[3 * i + 1], pz = xi[3 * i + 2];
        const double qx = vi[3 * i], qy = vi[3 * i + 1], qz = vi[3 * i + 2];
        double ax_ = 0, ay_ = 0, az_ = 0;
        if (order >= ORDER_H6) { ax_ = ai[3 * i]; ay_ = ai[3 * i + 1]; az_ = ai[3 * i + 2]; }
        const double h2 = h2i ? h2i[i] : -1.0;
        double a0 = 0, a1 = 0, a2 = 0, ph = 0, j0 = 0, j1 = 0, j2 = 0, s0 = 0, s1 = 0, s2 = 0;
        double dsmin = INFINITY;
        int nn = -1, cnt = 0, stored = 0;
        int *lst = (ngb_list && cap > 0) ? ngb_list + (size_t)i * cap : NULL;
        for (int j = 0; j < nj; ++j) {
            if (idi[i] == idj[j]) continue;                      /* kernels.cu:565 */
            const double dx = xj[3 * j] - px;
            const double dy = xj[3 * j + 1] - py;
            const double dz = xj[3 * j + 2] - pz;
            const double ds2 = fma(dz, dz, fma(dy, dy, dx * dx));
            if (ds2 < dsmin) { dsmin = ds2; nn = j; }            /* kernels.cu:541 */
            if (h2 > ds2) {                                      /* kernels.cu:584 */
                ++cnt;
                if (lst) keep_smallest(lst, &stored, cap, idj[j]);
            }
            const double inv = 1.0 / sqrt(ds2 + eps2);           /* hostclass.cpp:1408 */
            const double m = mj[j];
            const double minvr1 = m * inv;
            const double invr2 = inv * inv;
            const double minvr3 = minvr1 * invr2;
            a0 += minvr3 * dx;
            a1 += minvr3 * dy;
            a2 += minvr3 * dz;
            ph += (-1.0) * minvr1;
            if (order == ORDER_G5) continue;
            const double vx = vj[3 * j] - qx, vy = vj[3 * j + 1] - qy, vz = vj[3 * j + 2] - qz;
            const double rv = dx * vx + dy * vy + dz * vz;
            if (order == ORDER_H4) {
                const double drdv = (-3.0) * (minvr3 * invr2) * rv;  /* hostclass.cpp:1423 */
                j0 += minvr3 * vx + drdv * dx;
                j1 += minvr3 * vy + drdv * dy;
                j2 += minvr3 * vz + drdv * dz;
                continue;
            }
            /* SIXTH: kernels.cu:642-667 */
            const double ux = aj[3 * j] - ax_, uy = aj[3 * j + 1] - ay_, uz = aj[3 * j + 2] - az_;
            const double v2 = vx * vx + vy * vy + vz * vz;
            const double ra = dx * ux + dy * uy + dz * uz;
            double alpha = rv * invr2;
            double beta = (v2 + ra) * invr2 + alpha * alpha;
            alpha *= -3.0;
            const double jx = minvr3 * vx + alpha * (minvr3 * dx);
            const double jy = minvr3 * vy + alpha * (minvr3 * dy);
            const double jz = minvr3 * vz + alpha * (minvr3 * dz);
            alpha *= 2.0;
            beta *= -3.0;
            s0 += minvr3 * ux + alpha * jx + beta * (minvr3 * dx);
            s1 += minvr3 * uy + alpha * jy + beta * (minvr3 * dy);
            s2 += minvr3 * uz + alpha * jz + beta * (minvr3 * dz);
            j0 += jx;
            j1 += jy;
            j2 += jz;
        }
        acc[3 * i] = a0; acc[3 * i + 1] = a1; acc[3 * i + 2] = a2;
        pot[i] = ph;
        if (jrk) { jrk[3 * i] = j0; jrk[3 * i + 1] = j1; jrk[3 * i + 2] = j2; }
        if (snp && order >= ORDER_H6) { snp[3 * i] = s0; snp[3 * i + 1] = s1; snp[3 * i + 2] = s2; }
        if (nn_addr) nn_addr[i] = nn;
        if (nn_ds2) nn_ds2[i] = dsmin;
        if (ngb_cnt) ngb_cnt[i] = cnt;
        if (lst) qsort(lst, (size_t)stored, sizeof(int), cmp_int);
    }
}

/* Double-single split of a double (reference DSX::operator=(double),
 * kernels.cu:163-167): hi is x rounded to float, lo the rounded remainder. */
static inline void ds_split(double x, float *hi, float *lo)
{
    *hi = (float)x;
    *lo = (float)(x - (double)*hi);
}

/* Relative coordinate of the DS kernel (DSX::operator-, kernels.cu:170-178). */
static inline float ds_sub(float jh, float jl, float ih, float il)
{
    return (jh - ih) + (jl - il);
}

/* Nearest neighbour and neighbour lists with the mixed-precision (FP32)
 * distance contract of the header: the reference DS kernel keeps ds2 and
 * h2 in float (kernels.cu:419, 567-589).  nn_ds2 is the float ds2. */
void orc_neighbours_mixed(int ni, const double *xi, const double *h2i,
                          const int *idi, int nj, const double *xj,
                          const int *idj, int *nn_addr, float *nn_ds2,
                          int *ngb_cnt, int *ngb_list, int cap, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; ++i) {
        float ih[3], il[3];
        for (int d = 0; d < 3; ++d) ds_split(xi[3 * i + d], &ih[d], &il[d]);
        const float h2 = h2i ? (float)h2i[i] : -1.0f;
        float dsmin = INFINITY;
        int nn = -1, cnt = 0, stored = 0;
        int *lst = (ngb_list && cap > 0) ? ngb_list + (size_t)i * cap : NULL;
        for (int j = 0; j < nj; ++j) {
            if (idi[i] == idj[j]) continue;
            float jh[3], jl[3];
            for (int d = 0; d < 3; ++d) ds_split(xj[3 * j + d], &jh[d], &jl[d]);
            const float dx = ds_sub(jh[0], jl[0], ih[0], il[0]);
            const float dy = ds_sub(jh[1], jl[1], ih[1], il[1]);
            const float dz = ds_sub(jh[2], jl[2], ih[2], il[2]);
            const float ds2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (ds2 < dsmin) { dsmin = ds2; nn = j; }
            if (h2 > ds2) {
                ++cnt;
                if (lst) keep_smallest(lst, &stored, cap, idj[j]);
            }
        }
        if (nn_addr) nn_addr[i] = nn;
        if (nn_ds2) nn_ds2[i] = dsmin;
        if (ngb_cnt) ngb_cnt[i] = cnt;
        if (lst) qsort(lst, (size_t)stored, sizeof(int), cmp_int);
    }
}

/* Force-only FP64 direct summation (no neighbour bookkeeping): the timed CPU
 * baseline.  Same formulas as orc_force. */
void orc_force_plain(int order, int ni, const double *xi, const double *vi,
                     const double *ai, const int *idi, int nj, const double *xj,
                     const double *vj, const double *aj, const double *mj,
                     const int *idj, double eps2, double *acc, double *jrk,
                     double *snp, double *pot, int nthreads)
{
    orc_force(order, ni, xi, vi, ai, NULL, idi, nj, xj, vj, aj, mj, idj, eps2,
              acc, jrk, snp, pot, NULL, NULL, NULL, NULL, 0, nthreads);
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
