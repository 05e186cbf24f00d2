/* Thin C ABI between the C host driver and the HIP/CDNA4 kernels.
 *
 * The host driver (skg_host.c) never includes a HIP header: every device
 * allocation, copy, stream/graph operation and kernel launch goes through the
 * skk_* functions below, implemented in skg_kernels.hip.  All work is queued
 * on one library-owned stream (default priority).  Return codes: 0 ok,
 * otherwise a HIP error number (the text is in skk_error_text()).
 */
#ifndef SKG_KERNELS_H
#define SKG_KERNELS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ runtime */
int  skk_device_count(int *count);
int  skk_malloc(void **ptr, size_t bytes);
void skk_free(void *ptr);
int  skk_h2d(void *dst, const void *src, size_t bytes);        /* stream-ordered, waits */
int  skk_d2h(void *dst, const void *src, size_t bytes);        /* stream-ordered, waits */
int  skk_d2h_async(void *dst, const void *src, size_t bytes);  /* pinned dst only       */
int  skk_d2d(void *dst, const void *src, size_t bytes);        /* async                 */
int  skk_host_alloc(void **ptr, size_t bytes);                 /* pinned host memory    */
void skk_host_free(void *ptr);
int  skk_sync(void);
int  skk_last_launch_error(void);
void skk_reset_error(void);
const char *skk_error_text(int code);
/* Capture everything queued between begin and end into a graph. */
int  skk_graph_begin(void);
int  skk_graph_end(void **exec);
int  skk_graph_launch(void *exec);
void skk_graph_destroy(void *exec);
/* Event timing on the library stream (milliseconds between two marks). */
int    skk_mark(int slot);
double skk_elapsed_ms(int slot_a, int slot_b);
void  *skk_stream(void);

/* ------------------------------------------------------------- sparse */
#define SKK_MAX_BINS 12
#define SKK_BLOCK 256

/* Host-side row length classes that order the problem's rows and columns
 * (class k: rows of about 4 * 2^k entries, k 0..6; 8 above the long-row
 * threshold), so that rows of similar length are neighbours.  Bins own
 * consecutive row slots starting at row_begin. */
typedef struct skk_bin { int row_begin, row_end, log2L, block_begin; } skk_bin;
typedef struct skk_bins {
    int nbins, total_blocks;
    skk_bin b[SKK_MAX_BINS];
} skk_bins;

/* CSR-stream row tiles: tile t covers rows [tiles[2t], tiles[2t+2]) and
 * entries [tiles[2t+1], tiles[2t+3]),
 * consecutive rows holding at most a per-matrix entry cap together (a tile
 * of one row longer than SKK_TILE_NNZ is reduced by the whole workgroup)
 * and at most SKK_TILE_ROWS rows.  The cap is SKK_TILE_NNZ (the LDS
 * staging size), halved down to SKK_TILE_MIN_NNZ while the matrix would
 * have fewer than SKK_MIN_TILES tiles.  A product launches
 * min(ntiles, SKK_SPMV_GRID) workgroups; workgroup g takes tiles g,
 * g + grid, ..., so block partials never exceed SKK_SPMV_GRID per slot. */
#ifndef SKK_TILE_NNZ
#define SKK_TILE_NNZ 2048
#endif
#ifndef SKK_TILE_MIN_NNZ
#define SKK_TILE_MIN_NNZ 256
#endif
#ifndef SKK_MIN_TILES
#define SKK_MIN_TILES 1024
#endif
#ifndef SKK_TILE_ROWS
#define SKK_TILE_ROWS 1024
#endif
#ifndef SKK_SPMV_GRID
#define SKK_SPMV_GRID 4096
#endif
/* The step decision is taken inside the next dual prox (every workgroup
 * reducing the probe partials itself) while the A^T product and the step
 * probe have at most SKK_FUSE_PARTIALS workgroups together; SKK_FUSE_PER
 * partials per thread and slot are held in registers for it. */
#ifndef SKK_FUSE_PARTIALS
#define SKK_FUSE_PARTIALS 1024
#endif
#ifndef SKK_FUSE_PER
#define SKK_FUSE_PER 4
#endif
/* A matrix whose FP64 entries (12 bytes each) exceed this many bytes is read
 * with non-temporal loads: it cannot stay in the caches between products,
 * so it should not evict the gathered vector.  Smaller matrices are read
 * normally and are found in the 256 MB Infinity Cache by the next product
 * (config 2: 16% more iterations per second than streaming them). */
#ifndef SKK_STREAM_BYTES
#define SKK_STREAM_BYTES (64ll << 20)
#endif
/* Largest LDS copy of x (bytes per workgroup) an LP's step probe can keep:
 * with the 8 KB product buffer, two workgroups fit a CU's 160 KB. */
#define SKK_XCACHE_MAX 65536

typedef struct skk_csr {
    int rows, cols, nnz;
    const int *rp, *ci;
    const float *vf;         /* FP32 values (may be NULL) */
    const double *vd;        /* FP64 values               */
    const int *tiles;        /* ntiles + 1 pairs (first row, first entry) */
    int ntiles, grid;
    int stream;              /* 1: non-temporal matrix loads */
} skk_csr;

/* y = A x in FP64 (plain product, public ABI). */
int skk_spmv_f64(const skk_csr *A, const double *x, double *y);
/* y = A x with the iteration's FP32 values, FP32 gathers and FP32 sums
 * (the mixed engine's product kernel); A->vf must be present. */
int skk_spmv_f32(const skk_csr *A, const float *x, float *y);
int skk_axpy_f64(int n, double alpha, const double *x, double *y);

/* --------------------------------------------------- PDHG device state */

/* Device-resident scalars of the restarted PDHG loop (FP64).  The step
 * accept/reject decision, the buffer rotation that replaces a commit copy,
 * and the accepted-step counters all live here, so a whole restart window
 * runs without the host. */
typedef struct skk_state {
    double eta, omega, eta_min, eta_max;
    double updates, accept_count, accepted_at_restart;
    double dx2, dy2, inter, eta_bar;
    double eta_fixed;        /* fixed-step path: eta = 0.9/||A||           */
    int    accepted;         /* last step accepted                          */
    int    pending;          /* accepted iterate not yet added to the sums  */
    int    cur;              /* X[cur], Y[cur] hold the iterate             */
    int    a0, a1, a2;       /* AX[a0] = A x, AX[a1] = A x_prev, AX[a2] cand */
    int    pad[2];
} skk_state;

/* Iterate storage in one precision (float for MIXED, double for FP64). */
typedef struct skk_iter {
    void *X[2], *Y[2], *AX[3];
    void *xsum, *ysum, *xr, *yr, *xavg, *yavg, *xbar;
    void *QX[2];             /* QX[cur] = Q X[cur] (convex QP only)          */
    int sum64;               /* xsum, ysum hold doubles (else the set's type) */
} skk_iter;

typedef struct skk_pdhg {
    int m, n;
    skk_csr A, AT;                           /* AT is CSR(A^T)              */
    skk_csr Q;                               /* full symmetric Hessian      */
    int has_q;
    const float *cf, *lf, *uf, *rlf, *ruf;   /* FP32 copies of the data     */
    const double *cd, *ld, *ud, *rld, *rud;  /* FP64 data                   */
    skk_iter it[2];                          /* [0] FP64, [1] FP32          */
    double *ax_d, *aty_d, *qx_d;             /* FP64 products for checks    */
    skk_state *st;
    skk_state *snap;                         /* copy of st taken by the step
                                                probe, read by the fused
                                                decision of the next launch  */
    int fuse_decision;                       /* decide inside the next dual
                                                prox (few block partials)    */
    double *part;                            /* block partials              */
    double *eval_out;                        /* [2][8] residual sums        */
    int part_capacity;                       /* blocks per partial slot     */
    int elem_grid;                           /* fixed grid for reductions   */
    int xcache_bytes;                        /* LP step probe: LDS bytes of the
                                                last (longest) columns of x per
                                                workgroup, <= SKK_XCACHE_MAX
                                                (0: none)                    */
    int probe_grid;                          /* workgroups of the step probe
                                                (A, or Q and A in one launch) */
    /* Anchored refinement of the FP32 set (MIXED): its iterates hold the
     * correction (dx, dy) to an FP64 anchor (hx, hy), x = hx + dx,
     * y = hy + dy.  The FP32 data then describe the correction problem:
     * cf = c + A^T hy + Q hx, lf/uf = l/u - hx, rlf/ruf = rl/ru - A hx,
     * ycf = hy (the dual prox centre), all formed in FP64 and rounded
     * once.  In exact arithmetic the iterates are those of the unshifted
     * problem; in FP32 the rounding error scales with the correction. */
    int anchored;
    double *hx, *hy;                         /* anchor, FP64                */
    double *hax, *hg, *hqx;                  /* A hx, c + A^T hy + Q hx, Q hx */
    float *ycf;                              /* FP32 copy of hy             */
    /* Normalized-duality-gap restart merit (LP): each evaluation also
     * writes the NDG of the evaluated point against the last restart
     * point to eval_out[16 + which], in FP64. */
    int ndg;
    double *nx, *ng, *ny, *nax;              /* x, c + A^T y, y, A x (FP64)  */
    double *ndg_s;                           /* search state, 8 doubles     */
    double *ndg_part;                        /* 32 x part_capacity partials */
} skk_pdhg;

enum { SKK_F64 = 0, SKK_F32 = 1 };

int skk_pdhg_init(const skk_pdhg *P, int prec, double eta0, double omega, double eta_min,
                  double eta_max, int warm_x, int warm_y);
int skk_pdhg_adaptive_iter(const skk_pdhg *P, int prec);
/* A run of `count` adaptive iterations ending with the step decision
 * applied.  With P->fuse_decision the decision of iteration k is taken
 * inside the dual prox of iteration k + 1 (three launches per iteration
 * instead of four) and a closing one-workgroup decision follows the last. */
int skk_pdhg_adaptive_run(const skk_pdhg *P, int prec, int count);
int skk_pdhg_fixed_iter(const skk_pdhg *P, int prec);
/* One of the three launches of an adaptive iteration: 0 dual prox,
 * 1 A^T y_cand + primal prox, 2 A x_cand + step decision. */
int skk_pdhg_adaptive_part(const skk_pdhg *P, int prec, int part);
/* Event timeline on the library stream: reserve, mark slot k, read the
 * milliseconds between two marked slots (waits for the later one). */
int    skk_timeline_reserve(int count);
int    skk_timeline_mark(int k);
double skk_timeline_ms(int a, int b);
/* Random-gather issue rate: count 4-byte gathers from a table of `table`
 * floats, mean of `reps` launches; *rate in gathers per second. */
int skk_gather_probe(int table, int count, int reps, double *rate);
int skk_pdhg_flush(const skk_pdhg *P, int prec);
/* which 0: current iterate, 1: running average (formed here). */
int skk_pdhg_eval(const skk_pdhg *P, int prec, int which, int adaptive, double fixed_terms);
int skk_pdhg_restart(const skk_pdhg *P, int prec, int use_average, int adaptive);
int skk_pdhg_set_omega(const skk_pdhg *P, double omega);
/* Copy the FP32 iterate into the FP64 set and continue from it (with an
 * anchor, the FP64 set receives anchor + correction). */
int skk_pdhg_promote(const skk_pdhg *P);
/* Anchored refinement: the anchor absorbs the FP32 set's current iterate
 * (a restart has just put the restart point there), the FP32 corrections
 * return to zero, and the correction problem's FP32 data are formed again
 * from FP64 products at the new anchor.  absorb = 0 keeps the anchor
 * (set-up from hx, hy as uploaded). */
int skk_pdhg_reanchor(const skk_pdhg *P, int absorb);
/* Slot order -> the caller's order: output position o takes slot
 * col_slot[o] (row_slot[o]) times col_scale[o] (row_scale[o]); device
 * arrays in the original indexing. */
typedef struct skk_unpermute {
    const int *col_slot, *row_slot;
    const double *col_scale, *row_scale;
} skk_unpermute;
/* Store the final iterate (or the average) as FP64 into x_out/y_out, in
 * the original order and unscaled through U. */
int skk_pdhg_export(const skk_pdhg *P, int prec, int use_average, const skk_unpermute *U, double *x_out,
                    double *y_out);
/* v[0..n) = value on the library stream. */
int skk_fill_f64(double *v, double value, int n);
/* ||A||_2 by power iteration on A'A in FP64; v, v2 length n, w length m.
 * Leaves ||A'A v||^2 of the last step in *out (device). */
int skk_power_norm(const skk_pdhg *P, double *v, double *w, double *v2, int iters, double *out);
/* Reset the adaptive-step state to a new eta (FP64 tail start). */
int skk_pdhg_set_eta(const skk_pdhg *P, double eta);

/* ------------------------------------------------------ certificates */

/* sk_verify (c/src/sk_core.c:394-510) on the device in FP64 over the
 * original model: out[0] c'x, out[1] dual-objective terms, out[2] x'Qx,
 * out[3] primal infeasibility, out[4] dual infeasibility, out[5] largest
 * complementarity product.  act[m], rc[n] are written.  part holds
 * 8 x grid doubles. */
typedef struct skk_verify_in {
    int m, n;
    skk_csr A, AT;
    const skk_csr *Q;       /* may be NULL (full symmetric CSR) */
    const double *c, *rl, *ru, *l, *u, *x, *y;
    double *act, *rc, *qx;
    double *part; double *out; int grid;
} skk_verify_in;
int skk_verify(const skk_verify_in *V);

/* Relative KKT triple of an LP from device vectors with no allocation: the
 * block partials live in a static device array of 2 x 4 x SKK_SPMV_GRID
 * doubles (looked up once per device with hipGetSymbolAddress); out (device,
 * 5 doubles) receives primal, dual, gap, primal and dual objective.  Calls
 * that overlap on one device share that array and race. */
int skk_relkkt_noalloc(const skk_csr *A, const skk_csr *AT, const double *c, const double *rl, const double *ru,
                       const double *l, const double *u, const double *x, const double *y, double *out);

/* PDLP relative KKT sums in FP64 (same slots as the loop's evaluation):
 * out[0] ||Ax-proj Ax||^2, [1] dual violation^2, [2] c'x (+x'Qx), [3] dual
 * objective, [4] x'Qx.  Bounds with |b| >= 1e30 are absent. */
int skk_relative_kkt(const skk_verify_in *V);


/* ------------------------------------------- plain PDHG (legacy ABI) */
typedef struct skk_legacy {
    int m, n;
    const skk_csr *A, *AT, *Q;             /* Q: full symmetric CSR or NULL */
    double *x, *xbar, *y, *act, *grad, *qg, *prev;
    const double *c, *l, *u, *rl, *ru, *tau_vec, *sigma_vec, *qd;
    double tau, sigma, theta;
    double *part; double *out; int grid;
} skk_legacy;
int skk_legacy_init(double *x, double *xbar, const double *l, const double *u, int n, double *y, int m);
int skk_legacy_iter(const skk_legacy *L);
/* out[0] objective, [1] non-finite count, [2] step, [3] scale, [4] row
 * violation, [5] projected KKT, [6] dual residual, [7] complementarity */
int skk_legacy_check(const skk_legacy *L, int have_prev);

#ifdef __cplusplus
}
#endif

#endif
