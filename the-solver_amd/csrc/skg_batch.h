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
    double tail_switch;    /* FP32 -> FP64 handoff merit (mixed precision)  */
    int max_iterations, restart_period, power_iterations;
    int mixed;             /* 1: FP32 phase then FP64 tail; 0: FP64 only    */
    int tail_stall_windows;
    int anchored;          /* FP32 phase holds corrections to an FP64 anchor
                              kept in x_out / y_out (no FP64 tail)          */
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

/* Per-instance result, and the state the FP64 tail resumes from. */
typedef struct skb_out {
    int iterations, restarts, converged;
    int handoff;           /* 1: the FP32 phase stopped for the FP64 tail  */
    int fp32_iterations, pad;
    double rel_primal, rel_dual, rel_gap, matrix_norm, final_step;
    double omega, updates;
} skb_out;

typedef struct skb_batch {
    int count, max_m, max_n, pad;
    const skb_inst *inst;
    const int *order;              /* queue position -> instance          */
    const int *a_rp, *t_rp;
    const unsigned short *a_ci, *t_ci;  /* local column indices (< 65536) */
    const float *a_v, *t_v;
    const double *a_vd, *t_vd;
    const float *c, *l, *u, *rl, *ru;
    const double *cd, *ld, *ud, *rld, *rud;
    double *x_out, *y_out;         /* scaled-space solutions              */
    skb_out *out;
    unsigned *queue;               /* work-queue ticket                   */
    skb_params params;
} skb_batch;

/* LDS of one workgroup: 7 column and 9 row vectors of 4- or 8-byte words
 * and the row pointers of A and A^T (max_m + max_n + 2 ints). */
size_t skb_lds_bytes(int max_m, int max_n, int word_bytes);
/* One phase over the whole batch: fp64 = 0 runs the FP32 phase from a cold
 * start; fp64 = 1 runs FP64 iterates -- from a cold start when
 * params.mixed is 0, else resuming every instance the FP32 phase handed
 * off (the others are skipped). */
int    skb_launch(const skb_batch *B, int block_threads, int grid, int fp64);
int    skb_max_lds(void);
int    skb_cu_count(void);
/* Workgroups of one phase's kernel a CU holds at once (-1 if unknown). */
int    skb_blocks_per_cu(int block_threads, int fp64, size_t lds_bytes);

#ifdef __cplusplus
}
#endif

#endif
