/* Batched PDHG: device-side layout shared by the C host driver
 * (skg_batch_host.c) and the persistent kernel (skg_batch.hip). */
#ifndef SKG_BATCH_H
#define SKG_BATCH_H

#include <stddef.h>

#include "skg_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skb_params {
    double eps, eta_min_factor, eta_max_factor;
    int max_iterations, restart_period, power_iterations, pad;
} skb_params;

/* One instance: offsets into the packed arrays. */
typedef struct skb_inst {
    int m, n;
    int row, col;          /* offsets of its row / column vectors          */
    int a_rp, t_rp;        /* offsets of its CSR(A) / CSR(A^T) row pointers */
    int nz;                /* offset of its nonzeros (same for A and A^T)  */
    int pad;
    double cost_norm, bound_norm;
} skb_inst;

typedef struct skb_out {
    int iterations, restarts, converged, pad;
    double rel_primal, rel_dual, rel_gap, matrix_norm, final_step;
} skb_out;

typedef struct skb_batch {
    int count, max_m, max_n, pad;
    const skb_inst *inst;
    const int *order;              /* queue position -> instance          */
    const int *a_rp, *a_ci, *t_rp, *t_ci;
    const float *a_v, *t_v;
    const double *a_vd, *t_vd;
    const float *c, *l, *u, *rl, *ru;
    const double *cd, *ld, *ud, *rld, *rud;
    double *x_out, *y_out;         /* scaled-space solutions              */
    skb_out *out;
    unsigned *queue;               /* work-queue ticket                   */
    skb_params params;
} skb_batch;

size_t skb_lds_bytes(int max_m, int max_n);
int    skb_launch(const skb_batch *B, int block_threads, int grid);
int    skb_max_lds(void);
int    skb_cu_count(void);

#ifdef __cplusplus
}
#endif

#endif
