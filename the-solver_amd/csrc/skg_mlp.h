/* MLP kernels (skg_mlp.hip) behind a thin C ABI for the host driver. */
#ifndef SKG_MLP_H
#define SKG_MLP_H

#include <stdint.h>

#include "skg_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skm_epilogue {
    const float *bias;       /* per output column, or NULL            */
    const float *mask;       /* zero where mask[r*ldm + c] <= 0       */
    int ldm;
    int relu;
    float *c32;              /* FP32 output (ldc) or NULL             */
    uint16_t *c16;           /* BF16 output (ldc) or NULL             */
    int ldc;
} skm_epilogue;

/* C = op(A) op(B) with BF16 operands and FP32 accumulation; splits > 1
 * splits K over workgroups and needs splitk_buf of splits * M * N floats. */
int skm_gemm(int M, int N, int K, const uint16_t *A, int lda, int trans_a, const uint16_t *B, int ldb,
             int trans_b, const skm_epilogue *ep, float *splitk_buf, int splits);
int skm_mse_grad(int count, const float *P, const float *T, float scale, float *dP, uint16_t *dP16,
                 double *part, double *loss_sum);
/* db = column sums of dZ; scratch holds max_slices * cols floats. */
int skm_colsum(int rows, int cols, const float *dZ, float *db, float *scratch, int max_slices);
/* Adam step; also writes the BF16 copy of each weight block q (offset
 * w_off[q], length w_len[q] in p) to w_dst[q]. */
int skm_adam(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
             float eps, int step, const long long *w_off, const long long *w_len, uint16_t *const *w_dst);
int skm_to_bf16(long long n, const float *src, uint16_t *dst);
/* dst (cols x rows, leading dimension ld_dst) = transpose of src (rows x cols). */
int skm_transpose_bf16(int rows, int cols, const uint16_t *src, int ld_src, uint16_t *dst, int ld_dst);
int skm_gather_std(int rows, int cols, const int *idx, const float *src, const float *mean,
                   const float *stdv, float *dst32, uint16_t *dst16);
int skm_destandardize(int rows, int cols, const float *p, const float *mean, const float *stdv, float *y);

#ifdef __cplusplus
}
#endif

#endif
