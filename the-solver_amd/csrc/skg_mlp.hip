// Learned warm start: a 3-layer MLP trained and evaluated on the GPU with
// hand-written BF16 MFMA kernels (FP32 accumulation, FP32 master weights).
//
// No reference counterpart: the reference always cold-starts PDHG
// (c/src/sk_solve.c:871, cuda/src/sankhya_cuda_lp.cu:205-213).
//
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N] with op = identity or
// transpose, 128 x 128 x 32 tiles, 256 threads = 4 wavefronts each owning a
// 64 x 64 quadrant computed with v_mfma_f32_32x32x16_bf16.  Both LDS tiles
// are stored K-contiguous ([row][k] for A, [col][k] for B) so every MFMA
// operand is one 16-byte LDS read.  Long-K products (the gradient through
// the 70,000-wide output layer) split K over blockIdx.z and finish in a
// reduction kernel that applies the same epilogue.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "skg_mlp.h"

namespace {

hipStream_t mstream() { return static_cast<hipStream_t>(skk_stream()); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32, PAD = 8, THREADS = 256;
constexpr int VR = BK / 8;   // 16-byte vectors along k in one row of a tile

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

struct Epi {
    const float *bias;      // per column, or null
    const float *mask;      // keep where mask[r*ldm + c] > 0, or null
    int ldm;
    int relu;
    float *c32;             // fp32 output (ldc) or null
    uint16_t *c16;          // bf16 output (ldc) or null
    int ldc;
};

__device__ __forceinline__ void epilogue(const Epi &e, int r, int c, float v)
{
    if (e.bias) v += e.bias[c];
    if (e.relu) v = fmaxf(v, 0.0f);
    if (e.mask && !(e.mask[(size_t)r * e.ldm + c] > 0.0f)) v = 0.0f;
    if (e.c32) e.c32[(size_t)r * e.ldc + c] = v;
    if (e.c16) e.c16[(size_t)r * e.ldc + c] = f2bf(v);
}

// One operand tile of ROWS x BK elements into LDS as [row][k] (K-contiguous,
// the MFMA operand layout).  Source element (r, k) lives at
//   src[r * ld + k]  when KC (K-contiguous source), else src[k * ld + r].
// Loads are 16-byte vectors of 8 elements where the whole vector is in range
// and aligned (vec), otherwise element by element with zero fill.  A vector
// stays packed in its registers until store(), which runs after the
// current tile's MFMAs: nothing reads it earlier, so the load is in flight
// while they run (unpacking it at load time would make every load wait at
// once).
template <bool KC> struct TileLoader {
    static constexpr int kVec = (BM * BK) / (8 * THREADS);   // vectors per thread
    uint4 w[kVec];
    __device__ void load(const uint16_t *__restrict__ src, int ld, int r0, int rows, int k0, int kend, bool vec)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
            // KC: vector along k of one row; else: vector along rows of one k
            const int r = KC ? r0 + t / VR + q * (THREADS / VR) : r0 + ((t & 15) << 3);
            const int k = KC ? k0 + (t % VR) * 8 : k0 + (t >> 4) + q * (THREADS / 16);
            const bool full = KC ? (r < rows && k + 8 <= kend) : (k < kend && r + 8 <= rows);
            if (vec && full) {
                w[q] = *reinterpret_cast<const uint4 *>(KC ? src + (size_t)r * ld + k : src + (size_t)k * ld + r);
            } else {
                uint32_t h[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int rr = KC ? r : r + e, kk = KC ? k + e : k;
                    h[e] = (rr < rows && kk < kend) ? src[KC ? (size_t)rr * ld + kk : (size_t)kk * ld + rr] : 0u;
                }
                w[q] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                  h[6] | (h[7] << 16));
            }
        }
    }
    __device__ void store(uint16_t (*lds)[BK + PAD]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
            if (KC) {
                const int r = t / VR + q * (THREADS / VR), k = (t % VR) * 8;
                *reinterpret_cast<uint4 *>(&lds[r][k]) = w[q];
            } else {
                const int r = (t & 15) << 3, k = (t >> 4) + q * (THREADS / 16);
                const uint32_t p[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
#pragma unroll
                for (int e = 0; e < 8; ++e) lds[r + e][k] = static_cast<uint16_t>(p[e >> 1] >> (16 * (e & 1)));
            }
        }
    }
};

// C[M x N] = op(A)[M x K] op(B)[K x N], op(A)[i][k] = TA ? A[k*lda + i] : A[i*lda + k],
// op(B)[k][j] = TB ? B[j*ldb + k] : B[k*ldb + j].  128 x 128 x 32 tiles,
// four wavefronts of 64 x 64, each a 2 x 2 grid of v_mfma_f32_32x32x16_bf16
// (the CDNA4 double-rate BF16 MFMA); LDS double-buffered with one barrier
// per K step.  With kchunk > 0 the z-slice covers k in [z*kchunk,
// (z+1)*kchunk) and writes raw sums to part[z][M][N].
template <bool TA, bool TB>
__global__ __launch_bounds__(THREADS) void k_gemm(int M, int N, int K, const uint16_t *__restrict__ A,
                                                  int lda, const uint16_t *__restrict__ B, int ldb,
                                                  Epi e, float *part, int kchunk, int vec)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN][BK + PAD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
    int kbeg = 0, kend = K;
    if (kchunk > 0) {
        kbeg = blockIdx.z * kchunk;
        kend = min(K, kbeg + kchunk);
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[bi][bj][r] = 0.0f;
    TileLoader<!TA> la;   // A is K-contiguous unless transposed
    TileLoader<TB> lb;    // B is K-contiguous only when transposed
    const bool v = vec != 0;
    la.load(A, lda, row0, M, kbeg, kend, v);
    lb.load(B, ldb, col0, N, kbeg, kend, v);
    la.store(As[0]);
    lb.store(Bs[0]);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) {
            la.load(A, lda, row0, M, k0 + BK, kend, v);
            lb.load(B, ldb, col0, N, k0 + BK, kend, v);
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int kq = ks + ((lane >> 5) << 3);
            bf16x8 a[2], b[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                a[q] = *reinterpret_cast<const bf16x8 *>(&As[buf][wm * 64 + q * 32 + (lane & 31)][kq]);
                b[q] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][wn * 64 + q * 32 + (lane & 31)][kq]);
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
        }
        if (more) {
            la.store(As[buf ^ 1]);
            lb.store(Bs[buf ^ 1]);
        }
        __syncthreads();
        buf ^= 1;
    }
    // D[i][j] of one 32 x 32 block: j = lane % 32, i = 8 * (r / 4) + 4 * (lane / 32) + r % 4
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int j = col0 + wn * 64 + bj * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = row0 + wm * 64 + bi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (i >= M || j >= N) continue;
                if (kchunk > 0) part[((size_t)blockIdx.z * M + i) * N + j] = acc[bi][bj][r];
                else epilogue(e, i, j, acc[bi][bj][r]);
            }
        }
}

__global__ void k_splitk_reduce(int M, int N, int splits, const float *part, Epi e)
{
    const size_t total = (size_t)M * N;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int z = 0; z < splits; ++z) s += part[(size_t)z * total + idx];
        epilogue(e, (int)(idx / N), (int)(idx % N), s);
    }
}

// dP = 2 (P - T) / (B * out); block partials of sum (P - T)^2.
__global__ void k_mse_grad(int count, const float *P, const float *T, float scale, float *dP,
                           uint16_t *dP16, double *part)
{
    double acc = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const float d = P[k] - T[k];
        acc += (double)d * d;
        const float g = 2.0f * d * scale;
        dP[k] = g;
        dP16[k] = f2bf(g);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double sh[THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

__global__ void k_sum_partials(const double *part, int count, double *out)
{
    __shared__ double sh[THREADS];
    double acc = 0.0;
    for (int b = threadIdx.x; b < count; b += blockDim.x) acc += part[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// db[c] = sum_r dZ[r][c] in two fixed-order stages: slice s of the rows
// (blockIdx.y) gives part[s][c], then the slices are added in order.
__global__ void k_colsum_part(int rows, int cols, int slices, const float *dZ, float *part)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    if (c >= cols) return;
    const int r0 = (int)((long long)rows * s / slices), r1 = (int)((long long)rows * (s + 1) / slices);
    float acc = 0.0f;
    for (int r = r0; r < r1; ++r) acc += dZ[(size_t)r * cols + c];
    part[(size_t)s * cols + c] = acc;
}

__global__ void k_colsum_fin(int cols, int slices, const float *part, float *db)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float acc = 0.0f;
    for (int s = 0; s < slices; ++s) acc += part[(size_t)s * cols + c];
    db[c] = acc;
}

// Adam on the flat FP32 parameter vector; the BF16 copies of the three
// weight blocks are refreshed in the same pass (no second read of p).
struct Bf16Copies {
    long long o[3], n[3];    // offset and length of W1, W2, W3 in p
    uint16_t *dst[3];
};

__global__ void k_adam(long long n, float *p, const float *g, float *m, float *v, float lr, float b1,
                       float b2, float eps, float c1, float c2, Bf16Copies w)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        const float pk = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        p[k] = pk;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (k >= w.o[q] && k < w.o[q] + w.n[q]) w.dst[q][k - w.o[q]] = f2bf(pk);
    }
}

// dst[c * ld_dst + r] = src[r * ld_src + c] for r < rows, c < cols: 64 x 64
// tiles through LDS so both the reads and the writes are row-contiguous;
// with even extents and leading dimensions (pairs) each lane moves two
// elements per access.
__global__ __launch_bounds__(THREADS) void k_transpose_bf16(int rows, int cols, const uint16_t *__restrict__ src,
                                                            int ld_src, uint16_t *__restrict__ dst, int ld_dst,
                                                            int pairs)
{
    __shared__ uint16_t tile[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    if (pairs) {
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 lanes, 2 elements each
        for (int k = ty; k < 64; k += 8) {
            const int r = r0 + k, c = c0 + 2 * tx;
            uint32_t w = 0;
            if (r < rows && c < cols) w = *reinterpret_cast<const uint32_t *>(src + (size_t)r * ld_src + c);
            tile[k][2 * tx] = (uint16_t)(w & 0xffffu);
            tile[k][2 * tx + 1] = (uint16_t)(w >> 16);
        }
        __syncthreads();
        for (int k = ty; k < 64; k += 8) {
            const int c = c0 + k, r = r0 + 2 * tx;
            if (c < cols && r < rows)
                *reinterpret_cast<uint32_t *>(dst + (size_t)c * ld_dst + r) =
                    (uint32_t)tile[2 * tx][k] | ((uint32_t)tile[2 * tx + 1][k] << 16);
        }
        return;
    }
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;    // 64 x 4
    for (int k = ty; k < 64; k += 4) {
        const int r = r0 + k, c = c0 + tx;
        tile[k][tx] = (r < rows && c < cols) ? src[(size_t)r * ld_src + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int c = c0 + k, r = r0 + tx;
        if (c < cols && r < rows) dst[(size_t)c * ld_dst + r] = tile[tx][k];
    }
}

__global__ void k_to_bf16(long long n, const float *src, uint16_t *dst)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = f2bf(src[k]);
}

// rows of src gathered by idx, standardised: dst[r][c] = (src[idx[r]][c] - mean[c]) / std[c]
__global__ void k_gather_std(int rows, int cols, const int *idx, const float *src, const float *mean,
                             const float *stdv, float *dst32, uint16_t *dst16)
{
    const size_t total = (size_t)rows * cols;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total;
         k += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / cols), c = (int)(k % cols);
        const float v = (src[(size_t)idx[r] * cols + c] - mean[c]) / stdv[c];
        if (dst32) dst32[k] = v;
        if (dst16) dst16[k] = f2bf(v);
    }
}

// y = mean + std * p
__global__ void k_destandardize(int rows, int cols, const float *p, const float *mean, const float *stdv,
                                float *y)
{
    const size_t total = (size_t)rows * cols;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total;
         k += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(k % cols);
        y[k] = mean[c] + stdv[c] * p[k];
    }
}

int grid1d(size_t n)
{
    size_t g = (n + THREADS - 1) / THREADS;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (int)g;
}

}  // namespace

extern "C" {

int skm_gemm(int M, int N, int K, const uint16_t *A, int lda, int trans_a, const uint16_t *B, int ldb,
             int trans_b, const skm_epilogue *ep, float *splitk_buf, int splits)
{
    Epi e{ep->bias, ep->mask, ep->ldm, ep->relu, ep->c32, ep->c16, ep->ldc};
    const int kchunk = splits > 1 ? ((K + splits - 1) / splits + BK - 1) / BK * BK : 0;
    const int zs = splits > 1 ? (K + kchunk - 1) / kchunk : 1;
    const dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, zs);
    // 16-byte operand vectors need 16-byte aligned bases and leading dimensions
    const int vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 8 == 0 && ldb % 8 == 0;
#define GO(ta, tb) hipLaunchKernelGGL((k_gemm<ta, tb>), grid, dim3(THREADS), 0, mstream(), M, N, K, A, lda, B, ldb, e, splitk_buf, kchunk, vec)
    if (!trans_a && !trans_b) GO(false, false);
    else if (!trans_a && trans_b) GO(false, true);
    else if (trans_a && !trans_b) GO(true, false);
    else GO(true, true);
#undef GO
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return (int)err;
    if (splits > 1) {
        hipLaunchKernelGGL(k_splitk_reduce, dim3(grid1d((size_t)M * N)), dim3(THREADS), 0, mstream(), M, N, zs,
                           splitk_buf, e);
        err = hipGetLastError();
    }
    return (int)err;
}

int skm_mse_grad(int count, const float *P, const float *T, float scale, float *dP, uint16_t *dP16,
                 double *part, double *loss_sum)
{
    const int g = grid1d((size_t)count) > 2048 ? 2048 : grid1d((size_t)count);
    hipLaunchKernelGGL(k_mse_grad, dim3(g), dim3(THREADS), 0, mstream(), count, P, T, scale, dP, dP16, part);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return (int)err;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(THREADS), 0, mstream(), part, g, loss_sum);
    return (int)hipGetLastError();
}

int skm_colsum(int rows, int cols, const float *dZ, float *db, float *scratch, int max_slices)
{
    int slices = rows / 8;
    if (slices > max_slices) slices = max_slices;
    if (slices < 1) slices = 1;
    const int gx = (cols + THREADS - 1) / THREADS;
    hipLaunchKernelGGL(k_colsum_part, dim3(gx, slices), dim3(THREADS), 0, mstream(), rows, cols, slices, dZ, scratch);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return (int)err;
    hipLaunchKernelGGL(k_colsum_fin, dim3(gx), dim3(THREADS), 0, mstream(), cols, slices, scratch, db);
    return (int)hipGetLastError();
}

int skm_adam(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
             int step, const long long *w_off, const long long *w_len, uint16_t *const *w_dst)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    Bf16Copies w;
    for (int q = 0; q < 3; ++q) { w.o[q] = w_off[q]; w.n[q] = w_len[q]; w.dst[q] = w_dst[q]; }
    hipLaunchKernelGGL(k_adam, dim3(grid1d((size_t)n)), dim3(THREADS), 0, mstream(), n, p, g, m, v, lr, b1, b2,
                       eps, c1, c2, w);
    return (int)hipGetLastError();
}

int skm_transpose_bf16(int rows, int cols, const uint16_t *src, int ld_src, uint16_t *dst, int ld_dst)
{
    if (rows <= 0 || cols <= 0) return 0;
    // pairs: both extents and leading dimensions even, 4-byte aligned bases
    const int pairs = rows % 2 == 0 && cols % 2 == 0 && ld_src % 2 == 0 && ld_dst % 2 == 0 &&
                      (uintptr_t)src % 4 == 0 && (uintptr_t)dst % 4 == 0;
    hipLaunchKernelGGL(k_transpose_bf16, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(THREADS), 0, mstream(),
                       rows, cols, src, ld_src, dst, ld_dst, pairs);
    return (int)hipGetLastError();
}

int skm_to_bf16(long long n, const float *src, uint16_t *dst)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1d((size_t)n)), dim3(THREADS), 0, mstream(), n, src, dst);
    return (int)hipGetLastError();
}

int skm_gather_std(int rows, int cols, const int *idx, const float *src, const float *mean, const float *stdv,
                   float *dst32, uint16_t *dst16)
{
    hipLaunchKernelGGL(k_gather_std, dim3(grid1d((size_t)rows * cols)), dim3(THREADS), 0, mstream(), rows, cols,
                       idx, src, mean, stdv, dst32, dst16);
    return (int)hipGetLastError();
}

int skm_destandardize(int rows, int cols, const float *p, const float *mean, const float *stdv, float *y)
{
    hipLaunchKernelGGL(k_destandardize, dim3(grid1d((size_t)rows * cols)), dim3(THREADS), 0, mstream(), rows,
                       cols, p, mean, stdv, y);
    return (int)hipGetLastError();
}

}  // extern "C"
