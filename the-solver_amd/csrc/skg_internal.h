/* Internal declarations shared by the C host files of sankhya_gpu. */
#ifndef SKG_INTERNAL_H
#define SKG_INTERNAL_H

#include "sankhya_gpu.h"
#include "skg_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

int  skg_fail(const char *fmt, ...);
int  skg_hip_fail(const char *what, int code);
void skg_clear_error(void);
double skg_now(void);
int  skg_have_device(void);

/* Host CSR (int32 offsets/indices, FP64 values). */
typedef struct hcsr {
    int rows, cols, nnz;
    int *p, *j;
    double *v;
} hcsr;

int  hcsr_alloc(hcsr *a, int rows, int cols, long long nnz);
void hcsr_free(hcsr *a);
int  hcsr_transpose(const hcsr *a, hcsr *t);
int  hcsr_from_csc(const sk_csc *c, hcsr *out);
int  hcsr_validate(int rows, int cols, int nnz, const int *rp, const int *ci, const double *v);

/* Row-bin thresholds: a row of length len is reduced by 2^k lanes with
 * k = ceil(log2(ceil(len / nnz_per_lane))) capped at 6, or by a whole
 * workgroup when len > long_row. */
typedef struct skg_tuning {
    int nnz_per_lane;
    int long_row;
} skg_tuning;

void skg_tuning_default(skg_tuning *t);
int  skg_lanes_log2(int len, const skg_tuning *t);
int  skg_plan_bins(const int *rp, int rows, const skg_tuning *t, int *order, skk_bins *bins);
int  skg_plan_bins_keyed(const int *rp, int rows, const skg_tuning *t, const int *key, int key_range,
                         int *order, skk_bins *bins);

/* Device CSR with its row tiles. */
typedef struct dcsr {
    int *rp, *ci, *tiles;
    double *vd;
    float *vf;
    skk_csr k;
} dcsr;

int  skg_tile_nnz(long long nnz);
int  skg_plan_tiles(const int *rp, int rows, int tile_nnz, int **tile_row);
int  dcsr_upload(dcsr *d, const hcsr *a, int with_f32);
int  dcsr_retile(dcsr *d, const int *rp, int tile_nnz, int grid_cap, int stream);
void dcsr_free(dcsr *d);

/* Ruiz equilibration (src/lp/Scaling.cpp:11-92) and Pock-Chambolle alpha=1
 * composed into (R, C) (Scaling.cpp:111-152); at = CSR(A^T). */
void skg_ruiz(const hcsr *a, const hcsr *at, int max_iters, double tol, double *R, double *C);
void skg_pock_chambolle_into(const hcsr *a, const hcsr *at, double *R, double *C);

/* Process-wide tuning (set by sk_gpu_set_tuning, read at plan time). */
extern skg_tuning g_skg_tuning;

#ifdef __cplusplus
}
#endif

#endif
