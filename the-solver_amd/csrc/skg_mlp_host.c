/* Host driver of the warm-start MLP (C11): parameter layout, seeded
 * initialisation, mini-batch training with Adam, batched inference.  All
 * arithmetic runs in the BF16-MFMA / FP32 kernels of skg_mlp.hip. */
#define _POSIX_C_SOURCE 200809L
#include "skg_internal.h"
#include "skg_mlp.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

struct sk_gpu_mlp {
    int in, hid, out;
    long long np;                /* parameter count                          */
    long long o_w1, o_b1, o_w2, o_b2, o_w3, o_b3;
    float *p, *g, *m, *v;        /* device: params, grads, Adam moments      */
    uint16_t *w1b, *w2b, *w3b;   /* device BF16 weight copies [in][out]      */
    uint16_t *w1t, *w2t, *w3t;   /* and transposed [out][in]: every GEMM reads
                                    both operands K-contiguous               */
    float *in_mean, *in_std, *out_mean, *out_std;
    int step;
    /* work buffers for batches of up to `cap` rows */
    int cap;
    int *idx;
    float *x32, *t32, *h1, *h2, *pred, *dp, *dz2, *dz1, *ybuf, *splitk;
    uint16_t *x16, *h1b, *h2b, *dp16, *dz2b, *dz1b;
    uint16_t *x16t, *h1t, *h2t, *dpt, *dz2t, *dz1t;   /* [cols][rows] copies */
    double *part, *loss;
    float *colsum;               /* COLSUM_SLICES * max(hid, out) floats     */
    double *loss_dst;            /* where backward leaves the loss sum (NULL: loss) */
    const int *gidx;             /* batch row indices on the device (NULL: idx)   */
    long long splitk_floats;
};

#define COLSUM_SLICES 32
/* K splits of the long-K gradient product dP W3' (K = outputs): enough
 * 128 x 128 tiles to cover the device several times over */
#define MAX_SPLITS 128

static uint64_t sm64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)(sm64(s) >> 11) * 0x1.0p-53; }

static void free_work(sk_gpu_mlp *M)
{
    skk_free(M->idx); skk_free(M->x32); skk_free(M->t32); skk_free(M->h1); skk_free(M->h2);
    skk_free(M->pred); skk_free(M->dp); skk_free(M->dz2); skk_free(M->dz1); skk_free(M->ybuf);
    skk_free(M->splitk); skk_free(M->x16); skk_free(M->h1b); skk_free(M->h2b); skk_free(M->dp16);
    skk_free(M->dz2b); skk_free(M->dz1b); skk_free(M->part); skk_free(M->loss); skk_free(M->colsum);
    skk_free(M->x16t); skk_free(M->h1t); skk_free(M->h2t); skk_free(M->dpt); skk_free(M->dz2t); skk_free(M->dz1t);
    M->x16t = M->h1t = M->h2t = M->dpt = M->dz2t = M->dz1t = NULL;
    M->idx = NULL; M->x32 = M->t32 = M->h1 = M->h2 = M->pred = M->dp = M->dz2 = M->dz1 = M->ybuf = NULL;
    M->splitk = NULL; M->x16 = M->h1b = M->h2b = M->dp16 = M->dz2b = M->dz1b = NULL;
    M->part = M->loss = NULL;
    M->colsum = NULL;
    M->cap = 0;
}

static int ensure_work(sk_gpu_mlp *M, int rows)
{
    const size_t r = (size_t)rows;
    int rc = 0;
    if (rows <= M->cap) return 0;
    free_work(M);
    M->splitk_floats = (long long)MAX_SPLITS * rows * M->hid;
#define A(p, bytes) if (!rc) rc = skk_malloc((void **)&(p), (bytes))
    A(M->idx, sizeof(int) * r);
    A(M->x32, sizeof(float) * r * M->in); A(M->x16, 2 * r * M->in);
    A(M->t32, sizeof(float) * r * M->out);
    A(M->h1, sizeof(float) * r * M->hid); A(M->h1b, 2 * r * M->hid);
    A(M->h2, sizeof(float) * r * M->hid); A(M->h2b, 2 * r * M->hid);
    A(M->pred, sizeof(float) * r * M->out); A(M->ybuf, sizeof(float) * r * M->out);
    A(M->dp, sizeof(float) * r * M->out); A(M->dp16, 2 * r * M->out);
    A(M->dz2, sizeof(float) * r * M->hid); A(M->dz2b, 2 * r * M->hid);
    A(M->dz1, sizeof(float) * r * M->hid); A(M->dz1b, 2 * r * M->hid);
    A(M->x16t, 2 * r * M->in); A(M->h1t, 2 * r * M->hid); A(M->h2t, 2 * r * M->hid);
    A(M->dpt, 2 * r * M->out); A(M->dz2t, 2 * r * M->hid); A(M->dz1t, 2 * r * M->hid);
    A(M->splitk, sizeof(float) * (size_t)M->splitk_floats);
    A(M->part, sizeof(double) * 2048); A(M->loss, sizeof(double));
    A(M->colsum, sizeof(float) * COLSUM_SLICES * (size_t)(M->out > M->hid ? M->out : M->hid));
#undef A
    if (rc) { free_work(M); return skg_hip_fail("MLP work buffers", rc); }
    M->cap = rows;
    return 0;
}

long long sk_gpu_mlp_param_count(const sk_gpu_mlp *M) { return M ? M->np : 0; }

void sk_gpu_mlp_destroy(sk_gpu_mlp *M)
{
    if (!M) return;
    free_work(M);
    skk_free(M->p); skk_free(M->g); skk_free(M->m); skk_free(M->v);
    skk_free(M->w1b); skk_free(M->w2b); skk_free(M->w3b);
    skk_free(M->w1t); skk_free(M->w2t); skk_free(M->w3t);
    skk_free(M->in_mean); skk_free(M->in_std); skk_free(M->out_mean); skk_free(M->out_std);
    free(M);
}

/* The [out][in] BF16 weight copies from the [in][out] ones. */
static int transpose_weights(sk_gpu_mlp *M)
{
    int rc;
    if ((rc = skm_transpose_bf16(M->in, M->hid, M->w1b, M->hid, M->w1t, M->in)) ||
        (rc = skm_transpose_bf16(M->hid, M->hid, M->w2b, M->hid, M->w2t, M->hid)) ||
        (rc = skm_transpose_bf16(M->hid, M->out, M->w3b, M->out, M->w3t, M->hid)))
        return rc;
    return 0;
}

static int refresh_bf16(sk_gpu_mlp *M)
{
    int rc;
    if ((rc = skm_to_bf16((long long)M->in * M->hid, M->p + M->o_w1, M->w1b)) ||
        (rc = skm_to_bf16((long long)M->hid * M->hid, M->p + M->o_w2, M->w2b)) ||
        (rc = skm_to_bf16((long long)M->hid * M->out, M->p + M->o_w3, M->w3b)) ||
        (rc = transpose_weights(M)))
        return skg_hip_fail("weight conversion", rc);
    return 0;
}

int sk_gpu_mlp_create(int inputs, int hidden, int outputs, unsigned long long seed, sk_gpu_mlp **out)
{
    sk_gpu_mlp *M;
    float *h;
    uint64_t s = seed;
    long long k;
    int rc = 0, j;
    skg_clear_error();
    if (!out || inputs <= 0 || hidden <= 0 || outputs <= 0) return skg_fail("invalid MLP shape");
    *out = NULL;
    if (skg_have_device()) return -1;
    M = (sk_gpu_mlp *)calloc(1, sizeof(*M));
    if (!M) return skg_fail("out of host memory");
    M->in = inputs; M->hid = hidden; M->out = outputs;
    M->o_w1 = 0;
    M->o_b1 = M->o_w1 + (long long)inputs * hidden;
    M->o_w2 = M->o_b1 + hidden;
    M->o_b2 = M->o_w2 + (long long)hidden * hidden;
    M->o_w3 = M->o_b2 + hidden;
    M->o_b3 = M->o_w3 + (long long)hidden * outputs;
    M->np = M->o_b3 + outputs;
    h = (float *)calloc((size_t)M->np, sizeof(float));
    if (!h) { free(M); return skg_fail("out of host memory"); }
    {   /* Glorot-uniform weights, zero biases */
        const long long off[3] = {M->o_w1, M->o_w2, M->o_w3};
        const int fin[3] = {inputs, hidden, hidden}, fout[3] = {hidden, hidden, outputs};
        int l;
        for (l = 0; l < 3; ++l) {
            const double a = sqrt(6.0 / (double)(fin[l] + fout[l]));
            for (k = 0; k < (long long)fin[l] * fout[l]; ++k) h[off[l] + k] = (float)((2.0 * u01(&s) - 1.0) * a);
        }
    }
#define A(p, bytes) if (!rc) rc = skk_malloc((void **)&(p), (bytes))
    A(M->p, sizeof(float) * (size_t)M->np); A(M->g, sizeof(float) * (size_t)M->np);
    A(M->m, sizeof(float) * (size_t)M->np); A(M->v, sizeof(float) * (size_t)M->np);
    A(M->w1b, 2 * (size_t)inputs * hidden); A(M->w2b, 2 * (size_t)hidden * hidden);
    A(M->w3b, 2 * (size_t)hidden * outputs);
    A(M->w1t, 2 * (size_t)inputs * hidden); A(M->w2t, 2 * (size_t)hidden * hidden);
    A(M->w3t, 2 * (size_t)hidden * outputs);
    A(M->in_mean, sizeof(float) * inputs); A(M->in_std, sizeof(float) * inputs);
    A(M->out_mean, sizeof(float) * outputs); A(M->out_std, sizeof(float) * outputs);
#undef A
    if (!rc) rc = skk_h2d(M->p, h, sizeof(float) * (size_t)M->np);
    if (!rc) {
        memset(h, 0, sizeof(float) * (size_t)M->np);
        rc = skk_h2d(M->m, h, sizeof(float) * (size_t)M->np);
        if (!rc) rc = skk_h2d(M->v, h, sizeof(float) * (size_t)M->np);
        if (!rc) rc = skk_h2d(M->in_mean, h, sizeof(float) * inputs);
        if (!rc) rc = skk_h2d(M->out_mean, h, sizeof(float) * outputs);
        for (j = 0; j < (inputs > outputs ? inputs : outputs); ++j) h[j] = 1.0f;
        if (!rc) rc = skk_h2d(M->in_std, h, sizeof(float) * inputs);
        if (!rc) rc = skk_h2d(M->out_std, h, sizeof(float) * outputs);
    }
    free(h);
    if (rc) { sk_gpu_mlp_destroy(M); return skg_hip_fail("MLP allocation", rc); }
    if (refresh_bf16(M)) { sk_gpu_mlp_destroy(M); return -1; }
    *out = M;
    return 0;
}

int sk_gpu_mlp_set_scaling(sk_gpu_mlp *M, const float *in_mean, const float *in_std, const float *out_mean,
                           const float *out_std)
{
    int rc = 0;
    skg_clear_error();
    if (!M) return skg_fail("null MLP");
    if (in_mean) rc = skk_h2d(M->in_mean, in_mean, sizeof(float) * M->in);
    if (!rc && in_std) rc = skk_h2d(M->in_std, in_std, sizeof(float) * M->in);
    if (!rc && out_mean) rc = skk_h2d(M->out_mean, out_mean, sizeof(float) * M->out);
    if (!rc && out_std) rc = skk_h2d(M->out_std, out_std, sizeof(float) * M->out);
    return rc ? skg_hip_fail("MLP scaling upload", rc) : 0;
}

int sk_gpu_mlp_get_params(sk_gpu_mlp *M, float *params)
{
    int rc;
    if (!M || !params) return skg_fail("null argument");
    rc = skk_d2h(params, M->p, sizeof(float) * (size_t)M->np);
    return rc ? skg_hip_fail("params read", rc) : 0;
}

int sk_gpu_mlp_set_params(sk_gpu_mlp *M, const float *params)
{
    int rc;
    if (!M || !params) return skg_fail("null argument");
    if ((rc = skk_h2d(M->p, params, sizeof(float) * (size_t)M->np))) return skg_hip_fail("params write", rc);
    return refresh_bf16(M);
}

/* Forward pass on rows gathered from dev_x by M->idx (rows <= cap). */
static int forward(sk_gpu_mlp *M, int rows, const float *dev_x)
{
    skm_epilogue e;
    int rc;
    if ((rc = skm_gather_std(rows, M->in, M->gidx ? M->gidx : M->idx, dev_x, M->in_mean, M->in_std, NULL,
                             M->x16))) return rc;
    memset(&e, 0, sizeof(e));
    e.bias = M->p + M->o_b1; e.relu = 1; e.c32 = M->h1; e.c16 = M->h1b; e.ldc = M->hid;
    if ((rc = skm_gemm(rows, M->hid, M->in, M->x16, M->in, 0, M->w1t, M->in, 1, &e, NULL, 1))) return rc;
    e.bias = M->p + M->o_b2; e.c32 = M->h2; e.c16 = M->h2b;
    if ((rc = skm_gemm(rows, M->hid, M->hid, M->h1b, M->hid, 0, M->w2t, M->hid, 1, &e, NULL, 1))) return rc;
    memset(&e, 0, sizeof(e));
    e.bias = M->p + M->o_b3; e.c32 = M->pred; e.ldc = M->out;
    return skm_gemm(rows, M->out, M->hid, M->h2b, M->hid, 0, M->w3t, M->hid, 1, &e, NULL, 1);
}

/* Loss and gradients of one batch (targets gathered from dev_y). */
static int backward(sk_gpu_mlp *M, int rows, const float *dev_y)
{
    skm_epilogue e;
    int rc, splits;
    const float scale = 1.0f / ((float)rows * (float)M->out);
    if ((rc = skm_gather_std(rows, M->out, M->gidx ? M->gidx : M->idx, dev_y, M->out_mean, M->out_std, M->t32,
                             NULL))) return rc;
    if ((rc = skm_mse_grad(rows * M->out, M->pred, M->t32, scale, M->dp, M->dp16, M->part,
                           M->loss_dst ? M->loss_dst : M->loss))) return rc;
    memset(&e, 0, sizeof(e));
    /* dW3 = H2' dP, db3 = colsum dP */
    e.c32 = M->g + M->o_w3; e.ldc = M->out;
    if ((rc = skm_transpose_bf16(rows, M->hid, M->h2b, M->hid, M->h2t, rows)) ||
        (rc = skm_transpose_bf16(rows, M->out, M->dp16, M->out, M->dpt, rows)))
        return rc;
    if ((rc = skm_gemm(M->hid, M->out, rows, M->h2t, rows, 0, M->dpt, rows, 1, &e, NULL, 1))) return rc;
    if ((rc = skm_colsum(rows, M->out, M->dp, M->g + M->o_b3, M->colsum, COLSUM_SLICES))) return rc;
    /* dZ2 = (dP W3') * [H2 > 0]; K = out is long: split it */
    splits = (int)(M->splitk_floats / ((long long)rows * M->hid));
    if (splits > MAX_SPLITS) splits = MAX_SPLITS;
    if (splits < 1) splits = 1;
    memset(&e, 0, sizeof(e));
    e.mask = M->h2; e.ldm = M->hid; e.c32 = M->dz2; e.c16 = M->dz2b; e.ldc = M->hid;
    if ((rc = skm_gemm(rows, M->hid, M->out, M->dp16, M->out, 0, M->w3b, M->out, 1, &e, M->splitk, splits)))
        return rc;
    /* dW2 = H1' dZ2, db2 */
    memset(&e, 0, sizeof(e));
    e.c32 = M->g + M->o_w2; e.ldc = M->hid;
    if ((rc = skm_transpose_bf16(rows, M->hid, M->h1b, M->hid, M->h1t, rows)) ||
        (rc = skm_transpose_bf16(rows, M->hid, M->dz2b, M->hid, M->dz2t, rows)))
        return rc;
    if ((rc = skm_gemm(M->hid, M->hid, rows, M->h1t, rows, 0, M->dz2t, rows, 1, &e, NULL, 1))) return rc;
    if ((rc = skm_colsum(rows, M->hid, M->dz2, M->g + M->o_b2, M->colsum, COLSUM_SLICES))) return rc;
    /* dZ1 = (dZ2 W2') * [H1 > 0] */
    memset(&e, 0, sizeof(e));
    e.mask = M->h1; e.ldm = M->hid; e.c32 = M->dz1; e.c16 = M->dz1b; e.ldc = M->hid;
    if ((rc = skm_gemm(rows, M->hid, M->hid, M->dz2b, M->hid, 0, M->w2b, M->hid, 1, &e, NULL, 1))) return rc;
    /* dW1 = X' dZ1, db1 */
    memset(&e, 0, sizeof(e));
    e.c32 = M->g + M->o_w1; e.ldc = M->hid;
    if ((rc = skm_transpose_bf16(rows, M->in, M->x16, M->in, M->x16t, rows)) ||
        (rc = skm_transpose_bf16(rows, M->hid, M->dz1b, M->hid, M->dz1t, rows)))
        return rc;
    if ((rc = skm_gemm(M->in, M->hid, rows, M->x16t, rows, 0, M->dz1t, rows, 1, &e, NULL, 1))) return rc;
    return skm_colsum(rows, M->hid, M->dz1, M->g + M->o_b1, M->colsum, COLSUM_SLICES);
}

static int upload_rows(float **dst, const float *src, long long count)
{
    int rc = skk_malloc((void **)dst, sizeof(float) * (size_t)(count > 0 ? count : 1));
    if (!rc && count) rc = skk_h2d(*dst, src, sizeof(float) * (size_t)count);
    return rc;
}

int sk_gpu_mlp_gradients(sk_gpu_mlp *M, const float *X, const float *Y, int count, float *loss, float *grad)
{
    float *dx = NULL, *dy = NULL;
    int *idx = NULL, rc, k;
    double l = 0.0;
    skg_clear_error();
    if (!M || !X || !Y || count <= 0 || !grad) return skg_fail("invalid arguments");
    if (ensure_work(M, count)) return -1;
    idx = (int *)malloc(sizeof(int) * (size_t)count);
    if (!idx) return skg_fail("out of host memory");
    for (k = 0; k < count; ++k) idx[k] = k;
    rc = upload_rows(&dx, X, (long long)count * M->in);
    if (!rc) rc = upload_rows(&dy, Y, (long long)count * M->out);
    if (!rc) rc = skk_h2d(M->idx, idx, sizeof(int) * (size_t)count);
    if (!rc) rc = forward(M, count, dx);
    if (!rc) rc = backward(M, count, dy);
    if (!rc) rc = skk_d2h(&l, M->loss, sizeof(double));
    if (!rc) rc = skk_d2h(grad, M->g, sizeof(float) * (size_t)M->np);
    free(idx); skk_free(dx); skk_free(dy);
    if (rc) return skg_hip_fail("MLP gradients", rc);
    if (loss) *loss = (float)(l / ((double)count * M->out));
    return 0;
}

int sk_gpu_mlp_train(sk_gpu_mlp *M, const float *X, const float *Y, int count, int epochs, int batch, float lr,
                     unsigned long long seed, float *epoch_loss)
{
    float *dx = NULL, *dy = NULL;
    double *dloss = NULL, *hloss = NULL;
    int *perm = NULL, *dperm = NULL, rc = 0, ep, k, steps;
    uint64_t s = seed;
    const long long off[3] = {M ? M->o_w1 : 0, M ? M->o_w2 : 0, M ? M->o_w3 : 0};
    skg_clear_error();
    if (!M || !X || !Y || count <= 0 || epochs < 0 || batch <= 0 || !(lr > 0.0f)) return skg_fail("invalid arguments");
    if (batch > count) batch = count;
    if (ensure_work(M, batch)) return -1;
    steps = count / batch;
    perm = (int *)malloc(sizeof(int) * (size_t)count);
    hloss = (double *)malloc(sizeof(double) * (size_t)steps);
    if (!perm || !hloss) { free(perm); free(hloss); return skg_fail("out of host memory"); }
    rc = upload_rows(&dx, X, (long long)count * M->in);
    if (!rc) rc = upload_rows(&dy, Y, (long long)count * M->out);
    if (!rc) rc = skk_malloc((void **)&dloss, sizeof(double) * (size_t)steps);
    if (!rc) rc = skk_malloc((void **)&dperm, sizeof(int) * (size_t)count);
    for (k = 0; k < count; ++k) perm[k] = k;
    for (ep = 0; ep < epochs && !rc; ++ep) {
        double sum = 0.0;
        int b, nb = 0;
        for (k = count - 1; k > 0; --k) {       /* Fisher-Yates */
            const int j = (int)(u01(&s) * (double)(k + 1));
            const int t = perm[k]; perm[k] = perm[j]; perm[j] = t;
        }
        /* the epoch's order goes up once and the losses stay on the device
           until the end of the epoch: no host round trip per step */
        if (!rc) rc = skk_h2d(dperm, perm, sizeof(int) * (size_t)count);
        for (b = 0; b + batch <= count && !rc; b += batch, ++nb) {
            const long long len[3] = {(long long)M->in * M->hid, (long long)M->hid * M->hid,
                                      (long long)M->hid * M->out};
            uint16_t *const dst[3] = {M->w1b, M->w2b, M->w3b};
            M->gidx = dperm + b;
            M->loss_dst = dloss + nb;
            if ((rc = forward(M, batch, dx)) || (rc = backward(M, batch, dy))) break;
            ++M->step;
            if ((rc = skm_adam(M->np, M->p, M->g, M->m, M->v, lr, 0.9f, 0.999f, 1e-8f, M->step, off, len, dst)))
                break;
            if ((rc = transpose_weights(M))) break;
        }
        M->loss_dst = NULL;
        M->gidx = NULL;
        if (!rc && nb) rc = skk_d2h(hloss, dloss, sizeof(double) * (size_t)nb);
        for (k = 0; k < nb && !rc; ++k) sum += hloss[k] / ((double)batch * M->out);
        if (epoch_loss) epoch_loss[ep] = nb ? (float)(sum / nb) : 0.0f;
    }
    M->loss_dst = NULL;
    M->gidx = NULL;
    free(perm); free(hloss); skk_free(dx); skk_free(dy); skk_free(dloss); skk_free(dperm);
    return rc ? skg_hip_fail("MLP training", rc) : 0;
}

int sk_gpu_mlp_predict(sk_gpu_mlp *M, const float *X, int count, float *Y)
{
    float *dx = NULL;
    int *idx = NULL, rc = 0, b, k;
    const int chunk = 1024;
    skg_clear_error();
    if (!M || !X || !Y || count < 0) return skg_fail("invalid arguments");
    if (count == 0) return 0;
    if (ensure_work(M, count < chunk ? count : chunk)) return -1;
    idx = (int *)malloc(sizeof(int) * (size_t)chunk);
    if (!idx) return skg_fail("out of host memory");
    for (b = 0; b < count && !rc; b += chunk) {
        const int rows = count - b < chunk ? count - b : chunk;
        for (k = 0; k < rows; ++k) idx[k] = k;
        skk_free(dx);
        dx = NULL;
        if ((rc = upload_rows(&dx, X + (size_t)b * M->in, (long long)rows * M->in))) break;
        if ((rc = skk_h2d(M->idx, idx, sizeof(int) * (size_t)rows))) break;
        if ((rc = forward(M, rows, dx))) break;
        if ((rc = skm_destandardize(rows, M->out, M->pred, M->out_mean, M->out_std, M->ybuf))) break;
        rc = skk_d2h(Y + (size_t)b * M->out, M->ybuf, sizeof(float) * (size_t)rows * M->out);
    }
    free(idx); skk_free(dx);
    return rc ? skg_hip_fail("MLP inference", rc) : 0;
}
