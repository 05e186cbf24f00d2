// Batched restarted PDHG: thousands of small LPs in one persistent launch.
//
// One workgroup owns one instance at a time and takes the next one from an
// atomic work queue (instances ordered largest first, so the long ones start
// early and the tail is made of short ones).  The iterates of an instance
// live in LDS in FP32, with the row pointers of A and A^T; their entries
// (16-bit column indices, FP32 values for the iteration, FP64 values for
// the checks) stay in global memory, where a workgroup's few tens of KB
// remain L2-resident.  The whole PDLP control
// loop of src/cuda/GpuPdlp.cpp:252-534 -- ||A|| by power iteration, the
// adaptive step of PdlpKernels.cu:405-518, FP64 evaluation of the current
// and averaged iterate every `period` iterations, restarts (0.2 / 0.8 /
// 0.36) and the primal-weight update -- runs inside the workgroup with no
// host involvement.  Every reduction is a fixed-shape workgroup reduction,
// so an instance's result does not depend on which workgroup solved it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "skg_batch.h"
#include "skg_wave.h"

// Inner loops over a row's entries are unrolled by this factor: with rows
// of 5 to 10 entries the unrolled body keeps several matrix loads in flight
// per lane (one accumulator, so the summation order is unchanged).
// Measured on 4096 LPs with m in 20..400: 1.39x over no unrolling; 2 and 8
// are slower than 4.
#ifndef SKB_UNROLL
#define SKB_UNROLL 4
#endif

namespace {

hipStream_t bstream()
{
    return static_cast<hipStream_t>(skk_stream());
}

// Workgroup sum of S doubles; every thread receives the totals.  Without
// kTail the closing barrier is left to the caller, which must pass another
// barrier before the next wg_sum writes sh.
template <int BT, int S, bool kTail = true>
__device__ __forceinline__ void wg_sum(double (&v)[S], double *sh /* S * BT/64 */)
{
    constexpr int W = BT / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double t = wave_sum(v[s]);
        if (lane == 0) sh[s * W + wave] = t;
    }
    __syncthreads();
    // one sum at a time: unrolled over all S * W partials the compiler
    // issues every LDS read first and spills to hold them
#pragma unroll 1
    for (int s = 0; s < S; ++s) {
        double t = sh[s * W];
#pragma unroll
        for (int w = 1; w < W; ++w) t += sh[s * W + w];
        v[s] = t;
    }
    if (kTail) __syncthreads();
}

template <typename T> __device__ __forceinline__ T clampv(T v, T lo, T hi) { return fmin(fmax(v, lo), hi); }

// pow, exp and log out of line: inlined, their polynomial constants are
// hoisted into registers held for the whole kernel (about 40 VGPRs and
// many SGPRs), which limits how many workgroups share a CU.  They run
// once per 64 iterations or per restart.
__device__ __noinline__ double pow_ool(double x, double y) { return pow(x, y); }
__device__ __noinline__ double exp_ool(double x) { return exp(x); }
__device__ __noinline__ double log_ool(double x) { return log(x); }


// One instance's operators and data.  The operators' column indices are
// 16-bit (an instance's vectors must fit in LDS, so it has far fewer than
// 65536 rows or columns): 12 rather than 16 bytes per entry for each
// iteration to read, so the operators of the instances running on an XCD
// compete less for its L2.
struct Inst {
    int m, n;
    const int *arp, *trp;                  // CSR(A) and CSR(A^T) row pointers
    const uint16_t *aci, *tci;             // their local column indices
    const float *av, *tv;
    const double *avd, *tvd;
    const float *c, *l, *u, *rl, *ru;
    const double *cd, *ld, *ud, *rld, *rud;
    double cost_norm, bound_norm;
};

// The FP32 or the FP64 copy of an array, by the iterate precision T.
template <typename T>
__device__ __forceinline__ const T *pick(const float *f, const double *d)
{
    if constexpr (sizeof(T) == 4) return f;
    else return d;
}

__device__ Inst load_inst(const skb_batch &B, int k)
{
    Inst I;
    const skb_inst &d = B.inst[k];
    I.m = d.m; I.n = d.n;
    I.arp = B.a_rp + d.a_rp; I.aci = B.a_ci + d.nz; I.av = B.a_v + d.nz; I.avd = B.a_vd + d.nz;
    I.trp = B.t_rp + d.t_rp; I.tci = B.t_ci + d.nz; I.tv = B.t_v + d.nz; I.tvd = B.t_vd + d.nz;
    I.c = B.c + d.col; I.l = B.l + d.col; I.u = B.u + d.col;
    I.cd = B.cd + d.col; I.ld = B.ld + d.col; I.ud = B.ud + d.col;
    I.rl = B.rl + d.row; I.ru = B.ru + d.row; I.rld = B.rld + d.row; I.rud = B.rud + d.row;
    I.cost_norm = d.cost_norm; I.bound_norm = d.bound_norm;
    return I;
}

struct Kkt { double primal, dual, gap, pmove, dmove, worst; };

__device__ Kkt make_kkt(const double *s, double cost_norm, double bound_norm)
{
    Kkt k;
    k.primal = sqrt(fmax(0.0, s[0])) / (1.0 + bound_norm);
    k.dual = sqrt(fmax(0.0, s[1])) / (1.0 + cost_norm);
    k.gap = fabs(s[2] - s[3]) / (1.0 + fabs(s[2]) + fabs(s[3]));
    k.pmove = sqrt(fmax(0.0, s[6]));
    k.dmove = sqrt(fmax(0.0, s[7]));
    if (!isfinite(k.primal) || !isfinite(k.dual) || !isfinite(k.gap)) k.primal = k.dual = k.gap = INFINITY;
    k.worst = fmax(fmax(k.primal, k.dual), k.gap);
    return k;
}

// Eight residual sums of PdlpKernels.cu:101-209 in FP64 for the point
// (x * xs, y * ys) -- xs = ys = 1 for the iterate, 1/terms for the average.
// With an anchor (hx, hy: FP64, global) the vectors are corrections: the
// point is (hx + x * xs, hy + y * ys), and the restart point is the anchor.
template <int BT, typename T>
__device__ void residual_sums(const Inst &I, const T *x, const T *y, const T *xr, const T *yr,
                              double xs, double ys, const double *hx, const double *hy, double *sh,
                              double (&out)[8])
{
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < I.n; j += BT) {
        double aty = 0.0;
        const int p1 = I.trp[j + 1];
        if (hy) {
#pragma unroll SKB_UNROLL
            for (int p = I.trp[j]; p < p1; ++p) aty += I.tvd[p] * (hy[I.tci[p]] + ys * (double)y[I.tci[p]]);
        } else {
#pragma unroll SKB_UNROLL
            for (int p = I.trp[j]; p < p1; ++p) aty += I.tvd[p] * (ys * (double)y[I.tci[p]]);
        }
        const double xd = xs * (double)x[j];
        const double xj = hx ? hx[j] + xd : xd, lo = I.ld[j], hi = I.ud[j], r = I.cd[j] + aty;
        const double dx = hx ? xd : xj - (double)xr[j];
        acc[2] += I.cd[j] * xj;
        acc[4] += xj * xj;
        acc[6] += dx * dx;
        double viol = 0.0, term = 0.0;
        const bool lf = isfinite(lo), hf = isfinite(hi);
        if (lf && hf) term = fmin(r * lo, r * hi);
        else if (lf) { viol = fmax(0.0, -r); term = fmax(0.0, r) * lo; }
        else if (hf) { viol = fmax(0.0, r); term = fmin(0.0, r) * hi; }
        else viol = fabs(r);
        acc[1] += viol * viol;
        acc[3] += term;
    }
    for (int i = threadIdx.x; i < I.m; i += BT) {
        double ax = 0.0;
        const int p1 = I.arp[i + 1];
        if (hx) {
#pragma unroll SKB_UNROLL
            for (int p = I.arp[i]; p < p1; ++p) ax += I.avd[p] * (hx[I.aci[p]] + xs * (double)x[I.aci[p]]);
        } else {
#pragma unroll SKB_UNROLL
            for (int p = I.arp[i]; p < p1; ++p) ax += I.avd[p] * (xs * (double)x[I.aci[p]]);
        }
        const double rl = I.rld[i], ru = I.rud[i];
        const double pr = ax - fmin(fmax(ax, rl), ru);
        const double yd = ys * (double)y[i];
        const double yi = hy ? hy[i] + yd : yd, dy = hy ? yd : yi - (double)yr[i];
        acc[0] += pr * pr;
        acc[5] += yi * yi;
        acc[7] += dy * dy;
        double support = 0.0;
        if (yi > 0.0) support = isfinite(ru) ? yi * ru : 0.0;
        else if (yi < 0.0) support = isfinite(rl) ? yi * rl : 0.0;
        acc[3] -= support;
    }
    wg_sum<BT, 8>(acc, sh);
    for (int s = 0; s < 8; ++s) out[s] = acc[s];
}

// y = A x (LDS vectors), one thread per row.
template <int BT, typename T>
__device__ __forceinline__ void spmv_rows(int rows, const int *rp, const uint16_t *ci, const T *v, const T *x,
                                          T *y)
{
    for (int i = threadIdx.x; i < rows; i += BT) {
        T s = 0;
        const int p1 = rp[i + 1];
#pragma unroll SKB_UNROLL
        for (int p = rp[i]; p < p1; ++p) s += v[p] * x[ci[p]];
        y[i] = s;
    }
}

// One instance, start to finish (T = float: the FP32 phase, which may hand
// off to the FP64 tail; T = double: FP64 from a cold start, or the tail
// resuming from the FP32 phase's point and state).
template <int BT, typename T>
__device__ void solve_instance(const skb_batch &B, int k, T *lds, double *sh)
{
    const skb_params &P = B.params;
    constexpr bool kFp32 = sizeof(T) == 4;
    const bool resume = !kFp32 && P.mixed;
    skb_out &o = B.out[k];
    if (resume && !o.handoff) return;    // finished in the FP32 phase (uniform)
    Inst I = load_inst(B, k);
    const int m = I.m, n = I.n;
    const T *av = pick<T>(I.av, I.avd), *tv = pick<T>(I.tv, I.tvd);
    const T *cg = pick<T>(I.c, I.cd), *lg = pick<T>(I.l, I.ld), *ug = pick<T>(I.u, I.ud);
    const T *rlg = pick<T>(I.rl, I.rld), *rug = pick<T>(I.ru, I.rud);
    // LDS layout: 7 column vectors, 9 row vectors, then the row pointers
    T *X0 = lds, *X1 = X0 + n, *xsum = X1 + n, *xr = xsum + n, *cl = xr + n, *ll = cl + n, *ul = ll + n;
    T *Y0 = ul + n, *Y1 = Y0 + m, *AX0 = Y1 + m, *AX1 = AX0 + m, *AX2 = AX1 + m, *ysum = AX2 + m,
      *yr = ysum + m, *rll = yr + m, *rul = rll + m;
    for (int j = threadIdx.x; j < n; j += BT) { cl[j] = cg[j]; ll[j] = lg[j]; ul[j] = ug[j]; }
    for (int i = threadIdx.x; i < m; i += BT) { rll[i] = rlg[i]; rul[i] = rug[i]; }
    // the row pointers of A and A^T (m + n + 2 ints) follow in LDS: every
    // row of every product starts with them, and an LDS read is a few
    // hundred cycles shorter than an L2 hit
    {
        int *arp_l = reinterpret_cast<int *>(rul + m), *trp_l = arp_l + (m + 1);
        for (int i = threadIdx.x; i <= m; i += BT) arp_l[i] = I.arp[i];
        for (int j = threadIdx.x; j <= n; j += BT) trp_l[j] = I.trp[j];
        I.arp = arp_l;
        I.trp = trp_l;
    }
    __syncthreads();

    double anorm, eta, omega = 1.0, updates = 0.0;
    int iteration = 0, restarts = 0, fp32_iterations = 0;
    const skb_inst &d = B.inst[k];
    // Anchored FP32 phase: the iterates are corrections to an FP64 anchor
    // kept in the instance's output vectors, which absorbs the restart point
    // at every restart; xr is then always zero and yr holds the anchor's y
    // in FP32 as the dual prox centre.
    const bool anch = kFp32 && P.anchored;
    double *hx = anch ? B.x_out + d.col : nullptr;
    double *hy = anch ? B.y_out + d.row : nullptr;
    if (resume) {
        anorm = o.matrix_norm; eta = o.final_step; omega = o.omega; updates = o.updates;
        iteration = o.iterations; restarts = o.restarts; fp32_iterations = o.fp32_iterations;
    } else {
        // ||A||_2: power iteration on A'A from v = 1/sqrt(n) (GpuPdlp.cpp:179-211)
        double norm_sq = 0.0;
        const T v0 = (T)(1.0 / sqrt((double)(n > 0 ? n : 1)));
        for (int j = threadIdx.x; j < n; j += BT) X0[j] = v0;
        __syncthreads();
        for (int it = 0; it < P.power_iterations; ++it) {
            spmv_rows<BT, T>(m, I.arp, I.aci, av, X0, Y0);
            __syncthreads();
            double acc[1] = {0.0};
            for (int j = threadIdx.x; j < n; j += BT) {
                T s = 0;
                const int p1 = I.trp[j + 1];
#pragma unroll SKB_UNROLL
                for (int p = I.trp[j]; p < p1; ++p) s += tv[p] * Y0[I.tci[p]];
                X1[j] = s;
                acc[0] += (double)s * (double)s;
            }
            wg_sum<BT, 1>(acc, sh);
            norm_sq = acc[0];
            const double nrm = sqrt(norm_sq);
            const T inv = nrm > 0.0 ? (T)(1.0 / nrm) : (T)1;
            for (int j = threadIdx.x; j < n; j += BT) X0[j] = X1[j] * inv;
            __syncthreads();
        }
        anorm = sqrt(sqrt(fmax(0.0, norm_sq)));
        if (!(anorm > 0.0) || !isfinite(anorm)) anorm = 1.0;
        eta = 0.9 / anorm;
        if (I.cost_norm > 0.0 && I.bound_norm > 0.0) omega = fmin(1e6, fmax(1e-6, I.cost_norm / I.bound_norm));
    }
    const double eta_min = P.eta_min_factor / anorm, eta_max = P.eta_max_factor / anorm;
    double accepted = 0.0, accepted_at_restart = 0.0;

    if (resume) {
        // the FP64 tail starts a fresh restart window at the handed-off point
        for (int j = threadIdx.x; j < n; j += BT) {
            const T v = (T)B.x_out[d.col + j];
            X0[j] = v; X1[j] = v; xsum[j] = 0; xr[j] = v;
        }
        for (int i = threadIdx.x; i < m; i += BT) {
            const T v = (T)B.y_out[d.row + i];
            Y0[i] = v; Y1[i] = v; AX2[i] = 0; ysum[i] = 0; yr[i] = v;
        }
        __syncthreads();
        spmv_rows<BT, T>(m, I.arp, I.aci, av, X0, AX0);
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += BT) AX1[i] = AX0[i];
    } else {
        // cold start (GpuPdlp.cpp:278-291)
        for (int j = threadIdx.x; j < n; j += BT) { X0[j] = 0; X1[j] = 0; xsum[j] = 0; xr[j] = 0; }
        for (int i = threadIdx.x; i < m; i += BT) {
            Y0[i] = 0; Y1[i] = 0; AX0[i] = 0; AX1[i] = 0; AX2[i] = 0; ysum[i] = 0; yr[i] = 0;
        }
        if (anch) {
            for (int j = threadIdx.x; j < n; j += BT) hx[j] = 0.0;
            for (int i = threadIdx.x; i < m; i += BT) hy[i] = 0.0;
        }
    }
    __syncthreads();
    T *x = X0, *xc = X1, *y = Y0, *yc = Y1, *ax = AX0, *axp = AX1, *axc = AX2;

    double mu_at_restart = INFINITY, mu_prev = INFINITY, best_mu_seen = INFINITY;
    int since_restart = 0, converged = 0, final_avg = 0, handoff = 0, stall = 0;
    Kkt best{INFINITY, INFINITY, INFINITY, 0, 0, INFINITY};
    while (iteration < P.max_iterations) {
        const int period = min(P.restart_period, P.max_iterations - iteration);
        double g3 = 0.0, g6 = 0.0;
        // An accepted iterate joins the running sums inside the next
        // iteration's dual and primal loops, where each thread reads its
        // own entries of it anyway; the sums are only read across threads
        // at the window's end, after the flush and barrier below.
        bool pending = false;
        for (int t = 0; t < period; ++t) {
            if ((t & 63) == 0) {
                // the step rule's factors (k + 1)^-0.3 and (k + 1)^-0.6 for
                // the next 64 iterations, k = updates + 1 + lane: the same
                // pow calls as per iteration, 64 at a time
                const double kl = (updates + (double)(threadIdx.x & 63)) + 1.0;
                g3 = pow_ool(kl + 1.0, -0.3);
                g6 = pow_ool(kl + 1.0, -0.6);
            }
            const T sigma = (T)(eta * omega), tau = (T)(eta / omega);
            for (int i = threadIdx.x; i < m; i += BT) {
                const T yi = y[i];
                if (pending) ysum[i] += yi;
                const T v = yi + sigma * ((T)2 * ax[i] - axp[i]);
                // anchored: the prox written for the correction (centre -yr)
                yc[i] = anch ? fmin(fmax(-yr[i], v - sigma * rul[i]), v - sigma * rll[i])
                             : v - sigma * clampv<T>(v / sigma, rll[i], rul[i]);
            }
            __syncthreads();
            double acc[3] = {0.0, 0.0, 0.0};
            for (int j = threadIdx.x; j < n; j += BT) {
                T s = 0;
                const int p1 = I.trp[j + 1];
#pragma unroll SKB_UNROLL
                for (int p = I.trp[j]; p < p1; ++p) s += tv[p] * yc[I.tci[p]];
                const T xj = x[j];
                if (pending) xsum[j] += xj;
                const T xn = clampv<T>(xj - tau * (cl[j] + s), ll[j], ul[j]);
                xc[j] = xn;
                const double dd = (double)xn - (double)xj;
                acc[0] += dd * dd;
            }
            pending = false;
            __syncthreads();
            for (int i = threadIdx.x; i < m; i += BT) {
                T s = 0;
                const int p1 = I.arp[i + 1];
#pragma unroll SKB_UNROLL
                for (int p = I.arp[i]; p < p1; ++p) s += av[p] * xc[I.aci[p]];
                axc[i] = s;
                const double dy = (double)yc[i] - (double)y[i];
                acc[1] += dy * dy;
                acc[2] += dy * ((double)s - (double)ax[i]);
            }
            wg_sum<BT, 3, false>(acc, sh);   // two barriers pass before the next one
            const double inter = fabs(acc[2]);
            const double movement = 0.5 * (omega * acc[0] + acc[1] / omega);
            const double eta_bar = (inter > 0.0 && movement > 0.0) ? movement / inter : eta_max;
            const bool accept = eta <= eta_bar;
            const double kk = updates + 1.0;
            double next = fmin((1.0 - lane_value(g3, t & 63)) * eta_bar, (1.0 + lane_value(g6, t & 63)) * eta);
            if (!(next > 0.0) || !isfinite(next)) next = eta;
            eta = fmin(fmax(next, eta_min), eta_max);
            updates = kk;
            if (accept) {
                accepted += 1.0;
                T *t0 = x; x = xc; xc = t0;
                t0 = y; y = yc; yc = t0;
                t0 = axp; axp = ax; ax = axc; axc = t0;
                pending = true;
            }
        }
        if (pending) {
            for (int j = threadIdx.x; j < n; j += BT) xsum[j] += x[j];
            for (int i = threadIdx.x; i < m; i += BT) ysum[i] += y[i];
        }
        __syncthreads();
        iteration += period;
        since_restart += period;
        if (kFp32) fp32_iterations += period;
        // FP64 evaluation of the iterate and of the running average
        double sc[8], sa[8];
        residual_sums<BT, T>(I, x, y, xr, yr, 1.0, 1.0, hx, hy, sh, sc);
        double terms = accepted - accepted_at_restart;
        if (terms < 1.0) terms = 1.0;
        residual_sums<BT, T>(I, xsum, ysum, xr, yr, 1.0 / terms, 1.0 / terms, hx, hy, sh, sa);
        const Kkt cur = make_kkt(sc, I.cost_norm, I.bound_norm);
        const Kkt avg = make_kkt(sa, I.cost_norm, I.bound_norm);
        const bool avg_better = avg.worst < cur.worst;
        best = avg_better ? avg : cur;
        final_avg = avg_better;
        if (best.primal <= P.eps && best.dual <= P.eps && best.gap <= P.eps) { converged = 1; break; }
        const double mu = best.worst;
        if (kFp32 && P.mixed && !anch) {
            // hand off to FP64 iterates once the merit reaches tail_switch,
            // or when within 100x of it no window has improved the best
            // merit by 10% for tail_stall_windows windows
            if (mu < 0.9 * best_mu_seen) { best_mu_seen = mu; stall = 0; }
            else ++stall;
            if (iteration < P.max_iterations &&
                (mu <= P.tail_switch || (mu <= 100.0 * P.tail_switch && stall >= P.tail_stall_windows))) {
                handoff = 1;
                break;
            }
        }
        const bool sufficient = mu <= 0.2 * mu_at_restart;
        const bool necessary = mu <= 0.8 * mu_at_restart && mu > mu_prev;
        const bool artificial = (double)since_restart >= 0.36 * (double)iteration;
        mu_prev = mu;
        if (!(sufficient || necessary || artificial)) continue;
        // restart (GpuPdlp.cpp:453-498)
        const T inv = (T)(1.0 / terms);
        if (avg_better) {
            for (int j = threadIdx.x; j < n; j += BT) x[j] = xsum[j] * inv;
            for (int i = threadIdx.x; i < m; i += BT) y[i] = ysum[i] * inv;
            __syncthreads();
        }
        if (anch) {
            // the restart point becomes the anchor (x kept inside [l, u]);
            // the corrections return to zero and the correction problem's
            // FP32 data are formed from FP64 products at the new anchor
            for (int j = threadIdx.x; j < n; j += BT) {
                hx[j] = fmin(fmax(hx[j] + (double)x[j], I.ld[j]), I.ud[j]);
                x[j] = 0; xsum[j] = 0;
            }
            for (int i = threadIdx.x; i < m; i += BT) {
                hy[i] += (double)y[i];
                y[i] = 0; ysum[i] = 0; ax[i] = 0; axp[i] = 0;
            }
            __syncthreads();
            for (int j = threadIdx.x; j < n; j += BT) {
                double g = I.cd[j];
                const int p1 = I.trp[j + 1];
#pragma unroll SKB_UNROLL
                for (int p = I.trp[j]; p < p1; ++p) g += I.tvd[p] * hy[I.tci[p]];
                cl[j] = (T)g;
                ll[j] = __double2float_ru(I.ld[j] - hx[j]);
                ul[j] = __double2float_rd(I.ud[j] - hx[j]);
            }
            for (int i = threadIdx.x; i < m; i += BT) {
                double a = 0.0;
                const int p1 = I.arp[i + 1];
#pragma unroll SKB_UNROLL
                for (int p = I.arp[i]; p < p1; ++p) a += I.avd[p] * hx[I.aci[p]];
                rll[i] = (T)(I.rld[i] - a);
                rul[i] = (T)(I.rud[i] - a);
                yr[i] = (T)hy[i];
            }
            __syncthreads();
        } else {
            spmv_rows<BT, T>(m, I.arp, I.aci, av, x, ax);
            __syncthreads();
            for (int i = threadIdx.x; i < m; i += BT) { axp[i] = ax[i]; ysum[i] = 0; yr[i] = y[i]; }
            for (int j = threadIdx.x; j < n; j += BT) { xsum[j] = 0; xr[j] = x[j]; }
            __syncthreads();
        }
        accepted_at_restart = accepted;
        since_restart = 0;
        mu_at_restart = mu;
        mu_prev = INFINITY;
        ++restarts;
        final_avg = 0;
        if (best.pmove > 1e-14 && best.dmove > 1e-14) {
            const double upd = exp_ool(0.5 * log_ool(best.dmove / best.pmove) + 0.5 * log_ool(omega));
            if (isfinite(upd)) omega = fmin(1e6, fmax(1e-6, upd));
        }
    }
    // scaled-space point in FP64 (the host unscales; the tail resumes here)
    const double inv = 1.0 / fmax(1.0, accepted - accepted_at_restart);
    if (anch) {
        for (int j = threadIdx.x; j < n; j += BT)
            hx[j] = fmin(fmax(hx[j] + (final_avg ? (double)xsum[j] * inv : (double)x[j]), I.ld[j]), I.ud[j]);
        for (int i = threadIdx.x; i < m; i += BT) hy[i] += final_avg ? (double)ysum[i] * inv : (double)y[i];
    } else {
        for (int j = threadIdx.x; j < n; j += BT) B.x_out[d.col + j] = final_avg ? (double)xsum[j] * inv : (double)x[j];
        for (int i = threadIdx.x; i < m; i += BT) B.y_out[d.row + i] = final_avg ? (double)ysum[i] * inv : (double)y[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        o.iterations = iteration;
        o.restarts = restarts;
        o.converged = converged;
        o.handoff = handoff;
        o.fp32_iterations = fp32_iterations;
        o.rel_primal = best.primal;
        o.rel_dual = best.dual;
        o.rel_gap = best.gap;
        o.matrix_norm = anorm;
        o.final_step = eta;
        o.omega = omega;
        o.updates = updates;
    }
    __syncthreads();
}

// Register budget: at most 512 / SKB_WAVES_PER_EU VGPRs per lane, so that
// SKB_WAVES_PER_EU waves share a SIMD.  Unconstrained the kernel takes
// about 170 (two waves per SIMD); at 4 the compiler spills a few values of
// the once-per-window code, and 4096 LPs with m in 20..400 run 8% faster.
#ifndef SKB_WAVES_PER_EU
#define SKB_WAVES_PER_EU 4
#endif
template <int BT, typename T>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(SKB_WAVES_PER_EU))) void k_batch_pdlp(skb_batch B)
{
    extern __shared__ double lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    __shared__ double sh[8 * (BT / 64)];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0)
            next = (int)__hip_atomic_fetch_add(B.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        // uniform by construction: readfirstlane lets the compiler keep the
        // instance's pointers and sizes in scalar registers
        const int q = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (q >= B.count) break;       // every wave of the workgroup leaves together
        solve_instance<BT, T>(B, B.order[q], lds, sh);
    }
}

__global__ void k_queue_reset(unsigned *queue)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *queue = 0u;
}

template <typename T> void launch_phase(const skb_batch *B, int block_threads, int grid, size_t lds)
{
    switch (block_threads) {
    case 64:  hipLaunchKernelGGL((k_batch_pdlp<64, T>), dim3(grid), dim3(64), lds, bstream(), *B); break;
    case 128: hipLaunchKernelGGL((k_batch_pdlp<128, T>), dim3(grid), dim3(128), lds, bstream(), *B); break;
    case 256: hipLaunchKernelGGL((k_batch_pdlp<256, T>), dim3(grid), dim3(256), lds, bstream(), *B); break;
    case 512: hipLaunchKernelGGL((k_batch_pdlp<512, T>), dim3(grid), dim3(512), lds, bstream(), *B); break;
    case 1024: hipLaunchKernelGGL((k_batch_pdlp<1024, T>), dim3(grid), dim3(1024), lds, bstream(), *B); break;
    default: break;
    }
}

// Workgroups of k_batch_pdlp<block_threads, T> one CU holds at once with
// `lds` bytes of dynamic LDS (registers, LDS and wave slots), or -1.
template <typename T> int resident(int block_threads, size_t lds)
{
    int v = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (block_threads) {
    case 64:  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_batch_pdlp<64, T>, 64, lds); break;
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_batch_pdlp<128, T>, 128, lds); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_batch_pdlp<256, T>, 256, lds); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_batch_pdlp<512, T>, 512, lds); break;
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_batch_pdlp<1024, T>, 1024, lds); break;
    default: break;
    }
    return e == hipSuccess ? v : -1;
}

}  // namespace

extern "C" {

size_t skb_lds_bytes(int max_m, int max_n, int word_bytes)
{
    return (size_t)word_bytes * (7 * (size_t)max_n + 9 * (size_t)max_m) +
           sizeof(int) * ((size_t)max_m + (size_t)max_n + 2);
}

int skb_launch(const skb_batch *B, int block_threads, int grid, int fp64)
{
    if (block_threads != 64 && block_threads != 128 && block_threads != 256 && block_threads != 512 &&
        block_threads != 1024)
        return -1;
    const size_t lds = skb_lds_bytes(B->max_m, B->max_n, fp64 ? 8 : 4);
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, bstream(), B->queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (fp64) launch_phase<double>(B, block_threads, grid, lds);
    else launch_phase<float>(B, block_threads, grid, lds);
    return (int)hipGetLastError();
}

int skb_max_lds(void)
{
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return v;
}

int skb_blocks_per_cu(int block_threads, int fp64, size_t lds)
{
    return fp64 ? resident<double>(block_threads, lds) : resident<float>(block_threads, lds);
}

int skb_cu_count(void)
{
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
    return v;
}

}  // extern "C"
