/* Seeded synthetic workloads for the PDHG engine.
 *
 * One documented PRNG (SplitMix64 seeding xoshiro256**, uniforms from the top
 * 53 bits, normals by the Box-Muller cosine branch) so every host produces
 * the same instance bit for bit.  Models are returned in the reference's
 * canonical CSC form (sk_model, c/include/sankhya.h:82-98) with row indices
 * sorted inside each column; all arrays are malloc'd and released by
 * skw_model_free.
 */
#ifndef SK_WORKLOADS_H
#define SK_WORKLOADS_H

#include "sankhya_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

void skw_model_free(sk_model *m);

/* Bounded LP: nrow x ncol, exactly `per_col` distinct uniform rows per column
 * with N(0,1) values; x0 ~ U[0,1]^n; rlow = A x0 - U[0,1], rupp = A x0 + U[0,1]
 * (independent draws per row); bounds [0, 10]; c ~ U[-1, 1]. */
int skw_bounded_lp(sk_model *m, int nrow, int ncol, int per_col,
                   unsigned long long seed);

/* Power-law LP: column degrees from P(d=k) ~ (k+q)^-s on 1..max_degree with q
 * solved so that the mean is `mean_degree`; rows uniform and distinct inside
 * a column; values +-10^U[-3,3] with a random sign; a fraction `eq_fraction`
 * of rows (Bernoulli per row) are equalities rlow = rupp = A x0, the rest are
 * two-sided as in skw_bounded_lp; bounds [0,10]; c ~ U[-1,1].  *q_out gets q. */
int skw_zipf_lp(sk_model *m, int nrow, int ncol, double s, double mean_degree,
                int max_degree, double eq_fraction, unsigned long long seed,
                double *q_out);

/* Convex QP: A with `a_per_col` N(0,1) values per column; Q = D + B'B with D
 * diagonal U[0.1,10] and B b_rows x ncol with `b_per_row` N(0,1) values per
 * row at distinct columns; Q stored as the full symmetric CSC with duplicates
 * summed.  c ~ N(0,1), bounds [-5,5], x0 ~ U[-1,1], rows A x0 -+ U[0,1]. */
int skw_sparse_qp(sk_model *m, int nrow, int ncol, int a_per_col, int b_rows,
                  int b_per_row, unsigned long long seed);

/* Batch of bounded LPs: instance i has m_i ~ U{m_lo..m_hi}, n_i = 2 m_i,
 * `per_col` nonzeros per column, seed `seed * count + i`. */
int skw_batch_sizes(int count, int m_lo, int m_hi, unsigned long long seed,
                    int *m_out);

/* Parametric LP family: a fixed pattern with `per_col` N(0,1) values per
 * column, c0 ~ U[-1,1], x0 ~ U[0,1], b0 = A x0, P (ncol x dim) and
 * R (nrow x dim) N(0,1) * 0.1, row-major.  The model holds c0 and b0 -+ 0.5;
 * skw_parametric_instance gives theta's c = c0 + P theta and
 * b-+ = b0 -+ (0.5 + 0.1 |R theta|). */
int skw_parametric_family(sk_model *m, int nrow, int ncol, int per_col,
                          int dim, unsigned long long seed, double *P,
                          double *R, double *c0, double *b0);
void skw_parametric_instance(int nrow, int ncol, int dim, const double *P,
                             const double *R, const double *c0,
                             const double *b0, const double *theta, double *c,
                             double *rlow, double *rupp);
/* theta ~ N(0, I) draws, `count` x dim row-major. */
void skw_normal_matrix(int count, int dim, unsigned long long seed, double *out);

#ifdef __cplusplus
}
#endif

#endif
