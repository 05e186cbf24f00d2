// HIP/CDNA4 (gfx950) kernels and the thin C ABI the host driver calls.
//
// Design (MI355X-first, not a translation of the reference's CUDA):
//  * Sparse products are CSR gathers on two stored operators, A and A^T, so no
//    atomics are needed (the reference's transpose scatters with atomicAdd,
//    cuda/src/sankhya_cuda.cu:45-53).  Rows are cut into tiles of at most
//    2048 entries; a workgroup streams a tile's products into LDS with
//    coalesced loads and sums each row from LDS (CSR-stream), and a
//    power-law row longer than a tile gets a whole workgroup.  Matrix
//    streams use non-temporal loads so L2 keeps the gathered vector.
//  * The PDHG updates are epilogues of those products: the dual prox is
//    fused into A*x, the primal prox + extrapolation + running sums into
//    A^T*y (src/cuda/PdlpKernels.cu:71-97 are separate kernels there).
//  * Reductions are two-stage: wavefront butterfly + LDS inside a workgroup,
//    then a one-workgroup launch combines the per-block partials in a fixed
//    order, so every sum is bit-reproducible for a fixed grid (the property
//    PdlpKernels.cu:10-25 asks for).
//  * The adaptive-step accept/reject decision (PdlpKernels.cu:405-518) is
//    taken on the device and committed by rotating buffer indices in a
//    device-resident state struct instead of copying vectors, so a restart
//    window is a fixed sequence of launches that can be captured in a graph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <type_traits>

#include "skg_kernels.h"
#include "skg_wave.h"

namespace {

constexpr int kBlock = SKK_BLOCK;
static_assert(kBlock >= 128, "the decision kernels use a thread of the second wave");
constexpr int kWaves = kBlock / 64;
constexpr int kMaxDevices = 16;

hipStream_t g_stream[kMaxDevices] = {};
hipEvent_t g_marks[kMaxDevices][8] = {};
thread_local int g_last = 0;

hipStream_t stream()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
    if (!g_stream[dev]) {
        // A blocking stream (default flags, default priority): it orders with
        // the null stream both ways, so device-pointer operations keep the
        // reference's default-stream contract (cuda/include/sankhya_cuda.h:
        // 35-37) -- a caller's earlier null-stream copy into the input is
        // seen, and a later hipMemcpy of the output waits for the product.
        if (hipStreamCreate(&g_stream[dev]) != hipSuccess) g_stream[dev] = nullptr;
    }
    return g_stream[dev];
}

int check(hipError_t e)
{
    g_last = static_cast<int>(e);
    return static_cast<int>(e);
}

int launch_status()
{
    return check(hipGetLastError());
}

// ------------------------------------------------------------------ helpers

template <typename T> __device__ __forceinline__ T clampv(T v, T lo, T hi)
{
    // fmin/fmax: an infinite bound never binds
    return fmin(fmax(v, lo), hi);
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ float shfl_xor_d(float v, int m) { return __shfl_xor(v, m, 64); }

template <typename T> __device__ __forceinline__ T ntload(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// Matrix entry load: non-temporal for matrices too large to stay cached.
template <bool NT, typename T> __device__ __forceinline__ T mload(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}


// Reduce S doubles over the workgroup; thread s < S gets total s.
template <int S, bool kMax = false>
__device__ __forceinline__ void block_reduce(double (&v)[S], double (&out)[S])
{
    __shared__ double sh[S > 0 ? S : 1][kWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double t = kMax ? wave_max(v[s]) : wave_sum(v[s]);
        if (lane == 0) sh[s][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double t = sh[s][0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) t = kMax ? fmax(t, sh[s][w]) : t + sh[s][w];
        out[s] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ void store_coherent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_coherent(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block partials of S slots -> part[s * stride + blockIdx.x].  A later
// single-workgroup launch combines them: kernel boundaries order the writes,
// so no workgroup ever fences (an agent-scope fence on a multi-XCD part
// writes back and invalidates L2, which would evict the gathered vectors).
template <int S, bool kMax = false>
__device__ __forceinline__ void store_partials(double (&v)[S], double *part, int stride)
{
    double blk[S];
    block_reduce<S, kMax>(v, blk);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) part[(size_t)s * stride + blockIdx.x] = blk[s];
    }
}

// Combine `count` partials of one slot: strided per-thread sums in a fixed
// order, then the workgroup tree, so the total is bit-reproducible.
__device__ double combine(const double *part, int count, bool is_max)
{
    // four independent chains keep four loads in flight per thread
    const double z = is_max ? -INFINITY : 0.0;
    double a0 = z, a1 = z, a2 = z, a3 = z;
    int b = threadIdx.x;
    for (; b + 3 * kBlock < count; b += 4 * kBlock) {
        const double v0 = part[b], v1 = part[b + kBlock], v2 = part[b + 2 * kBlock], v3 = part[b + 3 * kBlock];
        if (is_max) { a0 = fmax(a0, v0); a1 = fmax(a1, v1); a2 = fmax(a2, v2); a3 = fmax(a3, v3); }
        else { a0 += v0; a1 += v1; a2 += v2; a3 += v3; }
    }
    for (; b < count; b += kBlock) a0 = is_max ? fmax(a0, part[b]) : a0 + part[b];
    double acc[1] = {is_max ? fmax(fmax(a0, a1), fmax(a2, a3)) : (a0 + a1) + (a2 + a3)};
    double out[1];
    if (is_max) block_reduce<1, true>(acc, out);
    else block_reduce<1>(acc, out);
    return out[0];
}

// Sums of S slots at once: the same fixed-order per-thread chains and
// workgroup tree as combine() (so the totals are bit-identical to it), but
// every partial of every slot is loaded up front, so the loads of all slots
// are in flight together instead of one slot's after another's.
// `mid` runs once every thread's partials have arrived, before the
// workgroup tree: work placed there waits for no load of its own.
struct NoMid {
    __device__ void operator()() const {}
};

template <int S, int kPer, class Mid = NoMid>
__device__ void combine_sums(const double *const (&part)[S], const int (&count)[S], double (&out)[S],
                             const Mid &mid = Mid{})
{
    const int tid = threadIdx.x;
    bool fits = true;
#pragma unroll
    for (int s = 0; s < S; ++s) fits = fits && count[s] <= kPer * kBlock;
    if (!fits) {
        mid();
#pragma unroll
        for (int s = 0; s < S; ++s) out[s] = combine(part[s], count[s], false);
        return;
    }
    // Branch-free loads (an index past the end reads the slot's first
    // partial, which is then not used): a conditional load per element
    // lets the compiler consume one slot's values before it requests the
    // next slot's, one memory round trip per slot.
    double v[S][kPer];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + k * kBlock;
            const double t = part[s][min(i, max(count[s] - 1, 0))];
            v[s][k] = i < count[s] ? t : 0.0;
        }
    __builtin_amdgcn_sched_barrier(0);   // every load requested before the first is used
    double acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int next = 0;   // first element the four-chain loop of combine() leaves
#pragma unroll
        for (int q = 0; q + 3 < kPer; q += 4) {
            if (tid + (q + 3) * kBlock < count[s]) {
                a0 += v[s][q]; a1 += v[s][q + 1]; a2 += v[s][q + 2]; a3 += v[s][q + 3];
                next = q + 4;
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (k >= next && tid + k * kBlock < count[s]) a0 += v[s][k];
        acc[s] = (a0 + a1) + (a2 + a3);
    }
    mid();
    block_reduce<S>(acc, out);
}

// out[s] for slots [0, nsum) (sums) then [nsum, nsum + nmax) (maxima).
__global__ __launch_bounds__(kBlock) void k_finalize(const double *part, int stride, int count,
                                                     int nsum, int nmax, double *out)
{
    for (int s = 0; s < nsum + nmax; ++s) {
        const double t = combine(part + (size_t)s * stride, count, s >= nsum);
        if (threadIdx.x == 0) out[s] = t;
    }
}

int finalize(const double *part, int stride, int count, int nsum, int nmax, double *out)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, stream(), part, stride, count, nsum, nmax, out);
    return launch_status();
}

// ----------------------------------------------------------- binned SpMV

// Row epilogues.  Every epilogue has row(r, sum, partials).  One with
// kPrefetch, load(r) -- the row's own inputs, which do not depend on the
// sum -- and finish(r, sum, inputs, partials) is driven in two halves: its
// loads are issued right after the tile's gathers, so they are in flight
// while the products go to LDS and the rows are summed.  finish(r, s,
// load(r), p) computes exactly what row(r, s, p) does.
struct NoIn {};
template <class E, std::enable_if_t<E::kPrefetch, int> = 0>
__device__ __forceinline__ auto epi_load(const E &e, int r, int) -> decltype(e.load(r))
{
    return e.load(r);
}
template <class E> __device__ __forceinline__ NoIn epi_load(const E &, int, long) { return NoIn{}; }
template <class E, typename AT, class In>
__device__ __forceinline__ auto epi_finish(const E &e, int r, AT s, const In &in, double *p, int)
    -> decltype(e.finish(r, s, in, p))
{
    e.finish(r, s, in, p);
}
template <class E, typename AT>
__device__ __forceinline__ void epi_finish(const E &e, int r, AT s, const NoIn &, double *p, long)
{
    e.row(r, s, p);
}

// Rows per lane when a tile has more rows than half the workgroup (a tile
// holds at most SKK_TILE_ROWS rows).
constexpr int kRowsPerLane = (SKK_TILE_ROWS + kBlock - 1) / kBlock;

// Lane-strided partial dot product of one row: the entries p, p + L,
// p + 2L, ... below e, the even ones of that sequence summed into s0 and
// the odd ones into s1, each in order.  Eight entries a round: their index
// and value loads, then their eight gathers, are issued back to back with
// no branch between them (a position past the end reads the row's last
// entry again and its product is not added), so a lane has eight gathers
// in flight -- a per-entry branch would make each gather wait for the
// previous one.
template <typename VT, typename XT, typename AT, bool NT = true, class XS = const XT *>
__device__ __forceinline__ AT row_part(const int *__restrict__ ci, const VT *__restrict__ val,
                                       const XS &x, int p, int e, int L)
{
    AT s0 = 0, s1 = 0;
    const int last = e - 1;
    for (; p < e; p += 8 * L) {
        int c[8];
        VT v[8];
        XT g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = min(p + u * L, last);
            c[u] = mload<NT>(ci + q);
            v[u] = mload<NT>(val + q);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (p + u * L < e) {
                const AT t = static_cast<AT>(v[u]) * static_cast<AT>(g[u]);
                if (u & 1) s1 += t;
                else s0 += t;
            }
        }
    }
    return s0 + s1;
}

// The products of a tile's entries [p0, p1) (p1 > p0) into prod, K entries
// per lane at positions p0 + k * kBlock + tid: all K index and value loads,
// then all K gathers, with no branch between them (positions past the end
// read the tile's last entry and are not stored), so the gathers are in
// flight together.  mid() runs after the gathers are issued and before
// their results are stored.
template <int K, typename VT, typename XT, typename AT, bool NT, class XS, class Mid>
__device__ __forceinline__ void tile_products(const int *__restrict__ ci, const VT *__restrict__ val,
                                              const XS &x, int p0, int p1, AT *prod, const Mid &mid)
{
    const int tid = threadIdx.x, last = p1 - 1;
    int c[K];
    VT v[K];
    XT g[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = min(p0 + k * kBlock + tid, last);
        c[k] = mload<NT>(ci + q);
        v[k] = mload<NT>(val + q);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) g[k] = x[c[k]];
    // mid()'s loads must stay behind the gathers, so that waiting for the
    // gathers below does not wait for them (vector-memory counters retire
    // in order).  Their addresses take an opaque zero produced here, which
    // the compiler cannot move its loads above, and nothing is scheduled
    // across this point.
    int zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero)::"memory");
    __builtin_amdgcn_sched_barrier(0);
    mid(zero);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = p0 + k * kBlock + tid;
        if (p < p1) prod[p - p0] = static_cast<AT>(v[k]) * static_cast<AT>(g[k]);
    }
}

// CSR-stream product over row tiles (the CSR-Adaptive scheme of Greathouse
// and Daga, SC'14): a tile's entries are multiplied by consecutive lanes --
// fully coalesced value and index loads -- into LDS, then each row of the
// tile is summed from LDS by 2^k lanes (k chosen from the tile's row count)
// or, when the tile has more rows than lanes, by one lane per row.  A tile
// of one row longer than SKK_TILE_NNZ is reduced by the whole workgroup
// straight from memory.  The gathered vector is x0, or -- when `sel` is
// given -- the candidate buffer of the rotating pair (x0, x1) named by the
// device state (cur ^ 1).  Sums are in a fixed order, so the result does
// not depend on the launch shape.
// One row tile [r0, r1) of a CSR product, summed into the epilogue.  A
// tile of one row longer than SKK_TILE_NNZ is reduced by the whole
// workgroup straight from memory; otherwise the tile's entries are
// multiplied by consecutive lanes into LDS (`prod`) and each row is summed
// from there.  Every thread of the workgroup calls this for the same tile.
// The tile's descriptor gives its rows [r0, r1) and entries [p0, p1)
// together, so its entries are requested one memory round trip after the
// descriptor rather than two.
template <typename VT, typename XT, typename AT, bool NT, class Epi, class XS>
__device__ __forceinline__ void spmv_tile(const int *__restrict__ rp, const int *__restrict__ ci,
                                          const VT *__restrict__ val, const XS &x, int r0, int r1, int p0,
                                          int p1, const Epi &e, double *part, AT *prod)
{
    using In = decltype(epi_load(e, 0, 0));
    constexpr int kPer = SKK_TILE_NNZ / kBlock;
    const int tid = threadIdx.x;
    const int nr = r1 - r0;
    if (p1 - p0 > SKK_TILE_NNZ) {
        // one long row: workgroup-strided partial sums, wave and LDS combine
        AT sum = row_part<VT, XT, AT, NT, XS>(ci, val, x, p0 + tid, p1, kBlock);
        sum = wave_sum(sum);
        if ((tid & 63) == 0) prod[tid >> 6] = sum;
        __syncthreads();
        if (tid == 0) {
            AT tot = prod[0];
            for (int w = 1; w < kWaves; ++w) tot += prod[w];
            e.row(r0, tot, part);
        }
        __syncthreads();
        return;
    }
    if (nr == 0) return;   // an empty matrix's one empty tile: no rows to finish
    // Reduction mapping: with nr * 2 <= kBlock, 2^lg lanes per row (lg =
    // floor(log2(kBlock / nr)) capped at 6) and lane 0 of the group
    // finishes it; otherwise one lane per row, rows tid, tid + kBlock, ...
    const bool grouped = nr * 2 <= kBlock;
    int lg = 0;
    if (grouped)
        while (lg < 6 && (nr << (lg + 1)) <= kBlock) ++lg;
    const int L = 1 << lg, g = tid >> lg, lane = tid & (L - 1);
    // The rows' offsets and epilogue inputs, loaded after the gathers (row
    // indices clamped to the tile, so the loads need no branch).
    int qa[kRowsPerLane], qb[kRowsPerLane];
    In in[kRowsPerLane];
    auto fetch = [&](int zero) {
        if (grouped) {
            const int row = r0 + min(g, nr - 1) + zero;
            qa[0] = rp[row] - p0;
            qb[0] = rp[row + 1] - p0;
            in[0] = epi_load(e, row, 0);
        } else {
#pragma unroll
            for (int k = 0; k < kRowsPerLane; ++k) {
                const int row = min(r0 + tid + k * kBlock, r1 - 1) + zero;
                qa[k] = rp[row] - p0;
                qb[k] = rp[row + 1] - p0;
                in[k] = epi_load(e, row, 0);
            }
        }
    };
    // products of the tile's entries, at most kPer per lane: 1, 2, 4 or 8
    // (the tile's entry count is the same for every lane)
    static_assert(kPer >= 1 && kPer <= 8, "tile_products covers up to 8 entries per lane");
    if (p1 > p0) {
        const int per = (p1 - p0 + kBlock - 1) / kBlock;
        if (per > 4) tile_products<8, VT, XT, AT, NT>(ci, val, x, p0, p1, prod, fetch);
        else if (per > 2) tile_products<4, VT, XT, AT, NT>(ci, val, x, p0, p1, prod, fetch);
        else if (per > 1) tile_products<2, VT, XT, AT, NT>(ci, val, x, p0, p1, prod, fetch);
        else tile_products<1, VT, XT, AT, NT>(ci, val, x, p0, p1, prod, fetch);
    } else {
        fetch(0);
    }
    __syncthreads();
    if (grouped) {
        AT sum = 0;
        if (g < nr)
            for (int q = qa[0] + lane; q < qb[0]; q += L) sum += prod[q];
        // the L lanes of a row combine through DPP swaps up to 16 lanes
        // (L is uniform; all lanes take part), then LDS permutes
        if (L >= 2) sum += dpp<0xB1>(sum);     // lane ^ 1
        if (L >= 4) sum += dpp<0x4E>(sum);     // lane ^ 2
        if (L >= 8) sum += dpp<0x141>(sum);    // the other four of the eight
        if (L >= 16) sum += dpp<0x140>(sum);   // the other eight of the row
        for (int off = 16; off < L; off <<= 1) sum += shfl_xor_d(sum, off);
        if (g < nr && lane == 0) epi_finish(e, r0 + g, sum, in[0], part, 0);
    } else {
#pragma unroll
        for (int k = 0; k < kRowsPerLane; ++k) {
            const int row = r0 + tid + k * kBlock;
            if (row < r1) {
                AT sum = 0;
                for (int q = qa[k]; q < qb[k]; ++q) sum += prod[q];
                epi_finish(e, row, sum, in[k], part, 0);
            }
        }
    }
    __syncthreads();   // prod is reused by the next tile
}

// Gathers of x[c] with the columns [c0, c0 + kc) served from an LDS copy:
// an LDS read for every lane (clamped) and a global load only for the lanes
// outside the copy.  The pointers carry their address spaces, so the two
// stay an LDS and a global load (a select between generic pointers would
// become one flat load).
template <typename XT> struct XCached {
    const __attribute__((address_space(1))) XT *x;
    const __attribute__((address_space(3))) XT *xs;
    int c0, kc;
    __device__ __forceinline__ XT operator[](int c) const
    {
        const unsigned o = static_cast<unsigned>(c - c0);
        const bool in = o < static_cast<unsigned>(kc);
        const XT lv = xs[in ? o : 0u];
        XT v = lv;
        if (!in) v = x[c];
        return v;
    }
};

// Copy n 16-byte chunks from global memory to LDS: eight loads per lane in
// flight, then their stores (an index past the end copies the last chunk
// again, so no load or store is conditional).
__device__ __forceinline__ void fill_lds(uint4 *dst, const uint4 *__restrict__ src, int n)
{
    for (int base = 0; base < n; base += 8 * kBlock) {
        uint4 t[8];
        int i[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            i[u] = min(base + u * kBlock + static_cast<int>(threadIdx.x), n - 1);
            t[u] = src[i[u]];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) dst[i[u]] = t[u];
    }
}

// The gathered vector is x1 when *cur is 0 and x0 when it is 1: `cur` points
// at the device state's buffer index (candidate x1 = X[cur ^ 1] of the pair
// x0 = X[0], x1 = X[1]), or at the tile table's first entry (always 0) with
// x1 = x0 when there is nothing to select.  The first tile's descriptor, the
// selector and the epilogue's state are requested before any of them is
// waited for (the grid never exceeds the tile count, so tile blockIdx.x
// exists): a workgroup's prologue costs one memory round trip, not three.
// With kXC each workgroup first copies x[c0, c0 + kc) into its dynamic LDS
// and gathers those columns from there.
template <typename VT, typename XT, typename AT, class Epi, bool NT, bool kXC = false>
__global__ __launch_bounds__(kBlock) void k_spmv(const int *__restrict__ rp,
                                                 const int *__restrict__ ci,
                                                 const VT *__restrict__ val, const XT *x0,
                                                 const XT *x1, const int *__restrict__ cur,
                                                 const int *__restrict__ tiles, int ntiles, Epi epi,
                                                 int c0, int kc)
{
    __shared__ AT prod[SKK_TILE_NNZ];
    int t = blockIdx.x;
    const int *d = tiles + 2 * t;
    int r0 = d[0], p0 = d[1], r1 = d[2], p1 = d[3];
    Epi e = epi;
    e.begin(epi);
    const XT *__restrict__ x = (*cur ^ 1) ? x1 : x0;
    double part[Epi::kSlots > 0 ? Epi::kSlots : 1] = {};
    auto run = [&](const auto &xs) {
        for (;;) {
            // the next tile's descriptor (clamped to the last) is in flight
            // while this tile is processed
            const int tn = t + gridDim.x;
            d = tiles + 2 * min(tn, ntiles - 1);
            const int n0 = d[0], n1 = d[1], n2 = d[2], n3 = d[3];
            spmv_tile<VT, XT, AT, NT>(rp, ci, val, xs, r0, r1, p0, p1, e, part, prod);
            if (tn >= ntiles) break;
            t = tn;
            r0 = n0; p0 = n1; r1 = n2; p1 = n3;
        }
    };
    if constexpr (kXC) {
        // the workgroup's copy of x[c0, c0 + kc) (kc * sizeof(XT) a
        // multiple of 16, c0 * sizeof(XT) too), then its tiles
        extern __shared__ uint4 xdyn[];
        fill_lds(xdyn, reinterpret_cast<const uint4 *>(x + c0), static_cast<int>(kc * sizeof(XT) / 16));
        __syncthreads();
        run(XCached<XT>{(const __attribute__((address_space(1))) XT *)x,
                        (const __attribute__((address_space(3))) XT *)xdyn, c0, kc});
    } else {
        run(x);
    }
    e.end(part);
}

// Two products of the same vector in one launch (the Hessian and A of a QP
// step probe): the tiles of the first matrix, then those of the second,
// dealt round robin over the workgroups; each epilogue keeps its own
// partials.
struct TileSet {
    const int *rp, *ci, *tiles;
    int ntiles;
};

template <typename VT, typename XT, typename AT, class Epi1, class Epi2, bool NT1, bool NT2>
__global__ __launch_bounds__(kBlock) void k_spmv_pair(TileSet m1, const VT *__restrict__ v1, TileSet m2,
                                                      const VT *__restrict__ v2, const XT *x0, const XT *x1,
                                                      const int *__restrict__ cur, Epi1 epi1, Epi2 epi2)
{
    __shared__ AT prod[SKK_TILE_NNZ];
    const int total = m1.ntiles + m2.ntiles;
    // as in k_spmv: the first descriptor, the selector and both states
    // requested together
    int t = blockIdx.x;
    const int *d = t < m1.ntiles ? m1.tiles + 2 * t : m2.tiles + 2 * (t - m1.ntiles);
    int r0 = d[0], p0 = d[1], r1 = d[2], p1 = d[3];
    const XT *__restrict__ x = (*cur ^ 1) ? x1 : x0;
    Epi1 e1 = epi1;
    e1.begin(epi1);
    Epi2 e2 = epi2;
    e2.begin(epi2);
    double part1[Epi1::kSlots > 0 ? Epi1::kSlots : 1] = {};
    double part2[Epi2::kSlots > 0 ? Epi2::kSlots : 1] = {};
    for (;;) {
        const int tn = t + gridDim.x, tc = min(tn, total - 1);
        d = tc < m1.ntiles ? m1.tiles + 2 * tc : m2.tiles + 2 * (tc - m1.ntiles);
        const int n0 = d[0], n1 = d[1], n2 = d[2], n3 = d[3];
        if (t < m1.ntiles) spmv_tile<VT, XT, AT, NT1>(m1.rp, m1.ci, v1, x, r0, r1, p0, p1, e1, part1, prod);
        else spmv_tile<VT, XT, AT, NT2>(m2.rp, m2.ci, v2, x, r0, r1, p0, p1, e2, part2, prod);
        if (tn >= total) break;
        t = tn;
        r0 = n0; p0 = n1; r1 = n2; p1 = n3;
    }
    e1.end(part1);
    e2.end(part2);
}

template <typename YT, typename AT> struct EpiStore {
    static constexpr int kSlots = 0;
    YT *y;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int r, AT s, double *) const { y[r] = static_cast<YT>(s); }
    __device__ void end(double *) const {}
};

template <typename VT, typename XT, typename AT, class Epi>
int launch_spmv(const skk_csr &A, const VT *vals, const XT *x0, const XT *x1,
                const skk_state *sel, const Epi &epi)
{
    if (A.ntiles == 0) return 0;
    // without a state to select by, the tile table's first entry (0) picks
    // x1, which is then x0
    const int *cur = sel ? &sel->cur : A.tiles;
    if (!sel) x1 = x0;
    if (A.stream)
        hipLaunchKernelGGL((k_spmv<VT, XT, AT, Epi, true>), dim3(A.grid), dim3(kBlock), 0,
                           stream(), A.rp, A.ci, vals, x0, x1, cur, A.tiles, A.ntiles, epi, 0, 0);
    else
        hipLaunchKernelGGL((k_spmv<VT, XT, AT, Epi, false>), dim3(A.grid), dim3(kBlock), 0,
                           stream(), A.rp, A.ci, vals, x0, x1, cur, A.tiles, A.ntiles, epi, 0, 0);
    return launch_status();
}

// ------------------------------------------------------------ PDHG state

// The running sums of accepted iterates (the averages' numerators) are held
// in the set's type or, with sum64, in FP64 (chosen per engine at set-up; a
// uniform branch at each use).
template <typename T> struct Sums {
    void *p;
    int wide;
    __device__ __forceinline__ void add(int k, T v) const
    {
        if (wide) static_cast<double *>(p)[k] += static_cast<double>(v);
        else static_cast<T *>(p)[k] += v;
    }
    __device__ __forceinline__ void zero(int k) const
    {
        if (wide) static_cast<double *>(p)[k] = 0.0;
        else static_cast<T *>(p)[k] = T(0);
    }
    __device__ __forceinline__ double get(int k) const
    {
        return wide ? static_cast<double *>(p)[k] : static_cast<double>(static_cast<T *>(p)[k]);
    }
};

template <typename T> struct Iter {
    T *X[2], *Y[2], *AX[3], *xr, *yr, *xavg, *yavg, *xbar, *QX[2];
    Sums<T> xsum, ysum;
    __host__ __device__ Iter() {}
    explicit Iter(const skk_iter &it)
    {
        for (int k = 0; k < 2; ++k) X[k] = static_cast<T *>(it.X[k]), Y[k] = static_cast<T *>(it.Y[k]);
        for (int k = 0; k < 3; ++k) AX[k] = static_cast<T *>(it.AX[k]);
        xsum = Sums<T>{it.xsum, it.sum64}; ysum = Sums<T>{it.ysum, it.sum64};
        xr = static_cast<T *>(it.xr); yr = static_cast<T *>(it.yr);
        xavg = static_cast<T *>(it.xavg); yavg = static_cast<T *>(it.yavg);
        xbar = static_cast<T *>(it.xbar);
        QX[0] = static_cast<T *>(it.QX[0]); QX[1] = static_cast<T *>(it.QX[1]);
    }
    // Buffer selection by a runtime index through selects, never by
    // indexing the pointer arrays (that would put them in scratch memory).
    __device__ __forceinline__ T *x(int k) const { return k ? X[1] : X[0]; }
    __device__ __forceinline__ T *y(int k) const { return k ? Y[1] : Y[0]; }
    __device__ __forceinline__ T *qx(int k) const { return k ? QX[1] : QX[0]; }
    __device__ __forceinline__ T *ax(int k) const { return k == 0 ? AX[0] : (k == 1 ? AX[1] : AX[2]); }
};

// yc: the dual prox centre of an anchored FP32 set (null otherwise).
template <typename T> struct Data {
    const T *c, *l, *u, *rl, *ru, *yc;
};

template <typename T> Data<T> data_of(const skk_pdhg &P);
template <> Data<float> data_of<float>(const skk_pdhg &P)
{
    return {P.cf, P.lf, P.uf, P.rlf, P.ruf, P.anchored ? P.ycf : nullptr};
}
template <> Data<double> data_of<double>(const skk_pdhg &P) { return {P.cd, P.ld, P.ud, P.rld, P.rud, nullptr}; }

// The dual prox of a PDHG step, y+ = v - sigma proj_[rl,ru](v / sigma)
// (PdlpKernels.cu:377-389).  With a centre yc (anchored set: y = yc + dy,
// v and the bounds already shifted by the anchor) the same map is written
// for the correction, dy+ = clamp(-yc, v - sigma ru, v - sigma rl): the
// three cases of the prox (below, between, above the bounds) give
// v - sigma rl, -yc and v - sigma ru, so no term of the anchor's size is
// added and then cancelled in FP32.  Infinite bounds never bind.
// The dual prox from row i's bounds (and, anchored, its centre yc).
template <typename T> __device__ __forceinline__ T dual_prox_of(T v, T sigma, T rl, T ru, T yc, bool anchored)
{
    if (anchored) return fmin(fmax(-yc, v - sigma * ru), v - sigma * rl);
    return v - sigma * clampv(v / sigma, rl, ru);
}

template <typename T> __device__ __forceinline__ T dual_prox(T v, T sigma, const Data<T> &D, int i)
{
    return dual_prox_of(v, sigma, D.rl[i], D.ru[i], D.yc ? D.yc[i] : T(0), D.yc != nullptr);
}

// Anchor of an FP32 set in evaluations and exports (null x: none).
struct Anchor {
    const double *x, *y, *ax, *g, *qx;
};

Anchor anchor_of(const skk_pdhg &P, bool f32)
{
    if (!f32 || !P.anchored) return Anchor{nullptr, nullptr, nullptr, nullptr, nullptr};
    return Anchor{P.hx, P.hy, P.hax, P.hg, P.has_q ? P.hqx : nullptr};
}

template <typename T> const T *vals_of(const skk_csr &A);
template <> const float *vals_of<float>(const skk_csr &A) { return A.vf; }
template <> const double *vals_of<double>(const skk_csr &A) { return A.vd; }

// Partial-slot regions (each part_capacity doubles).
enum { R_DX = 0, R_DY = 1, R_INTER = 2, R_EVAL = 3, R_POWER = 11, R_VERIFY = 12, R_DQ = 13, R_DXF = 14 };
enum { C_PROBE = 0, C_EVAL = 1, C_POWER = 2, C_VERIFY = 3, C_KKT = 4 };

__global__ void k_init_state(skk_state *st, double eta0, double omega, double eta_min,
                             double eta_max)
{
    if (threadIdx.x || blockIdx.x) return;
    st->eta = eta0; st->omega = omega; st->eta_min = eta_min; st->eta_max = eta_max;
    st->updates = 0.0; st->accept_count = 0.0; st->accepted_at_restart = 0.0;
    st->dx2 = st->dy2 = st->inter = 0.0; st->eta_bar = eta0; st->eta_fixed = eta0;
    st->accepted = 1; st->pending = 0; st->cur = 0; st->a0 = 0; st->a1 = 1; st->a2 = 2;
}

__global__ void k_set_omega(skk_state *st, double omega)
{
    if (threadIdx.x || blockIdx.x) return;
    st->omega = omega;
}

__global__ void k_set_eta(skk_state *st, double eta)
{
    if (threadIdx.x || blockIdx.x) return;
    st->eta = eta;
    st->eta_fixed = eta;
}

__global__ void k_clear_pending(skk_state *st)
{
    if (threadIdx.x || blockIdx.x) return;
    st->pending = 0;
}

// Cold/warm start: X[0] = x0 (already in place when warm) or 0, Y likewise,
// sums 0, restart points = start point, xbar = x, AX[0] = AX[1] = 0 unless
// warm (then filled by a following product).
template <typename T>
__global__ void k_init_vectors(Iter<T> I, int n, int m, int warm_x, int warm_y)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const T x = warm_x ? I.X[0][j] : T(0);
        I.X[0][j] = x; I.X[1][j] = x; I.xsum.zero(j); I.xr[j] = x; I.xbar[j] = x; I.xavg[j] = x;
        if (I.QX[0]) { I.QX[0][j] = T(0); I.QX[1][j] = T(0); }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const T y = warm_y ? I.Y[0][i] : T(0);
        I.Y[0][i] = y; I.Y[1][i] = y; I.ysum.zero(i); I.yr[i] = y; I.yavg[i] = y;
        I.AX[0][i] = T(0); I.AX[1][i] = T(0); I.AX[2][i] = T(0);
    }
}

template <typename T> __global__ void k_copy(T *dst, const T *src, int n)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) dst[j] = src[j];
}

// ------------------------------------------------------ adaptive PDHG step

// y_cand = prox(y + sigma (2 A x - A x_prev)), sigma = eta * omega
// (PdlpKernels.cu:377-389); adds a pending accepted y into the running sum.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dual_adaptive(Iter<T> I, Data<T> D,
                                                          const skk_state *__restrict__ st, int m)
{
    const int cur = st->cur, pending = st->pending;
    const T sigma = static_cast<T>(st->eta * st->omega);
    const T *__restrict__ y = I.y(cur);
    const T *__restrict__ ax = I.ax(st->a0);
    const T *__restrict__ axp = I.ax(st->a1);
    T *__restrict__ yc = I.y(cur ^ 1);
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const T yi = y[i];
        const T v = yi + sigma * (T(2) * ax[i] - axp[i]);
        yc[i] = dual_prox(v, sigma, D, i);
        if (pending) I.ysum.add(i, yi);
    }
}

// A^T y_cand epilogue: x_cand = proj(x - tau (c + A^T y_cand)),
// tau = eta / omega (PdlpKernels.cu:391-400); partial ||x_cand - x||^2.
template <typename T, typename AT> struct EpiPrimalAdaptive {
    static constexpr int kSlots = 1;
    Iter<T> I;
    Data<T> D;
    const skk_state *st;
    double *part;
    int stride;
    // per-launch values
    const T *x, *qx;
    T *xc;
    T tau;
    int pending;
    template <class K> __device__ void begin(const K &)
    {
        const int cur = st->cur;
        x = I.x(cur); xc = I.x(cur ^ 1);
        qx = I.QX[0] ? I.qx(cur) : nullptr;
        tau = static_cast<T>(st->eta / st->omega);
        pending = st->pending;
    }
    static constexpr bool kPrefetch = sizeof(T) == 4;
    struct In {
        T x, c, l, u, q;
    };
    __device__ In load(int j) const { return In{x[j], D.c[j], D.l[j], D.u[j], qx ? qx[j] : T(0)}; }
    __device__ void finish(int j, AT s, const In &v, double *p) const
    {
        const T g = v.c + static_cast<T>(s) + v.q;
        const T xn = clampv(v.x - tau * g, v.l, v.u);
        xc[j] = xn;
        if (pending) I.xsum.add(j, v.x);
        const double d = static_cast<double>(xn) - static_cast<double>(v.x);
        p[0] += d * d;
    }
    __device__ void row(int j, AT s, double *p) const { finish(j, s, load(j), p); }
    __device__ void end(double *p) const
    {
        double v[1] = {p[0]};
        store_partials<1>(v, part, 0);
    }
};

// A x_cand epilogue: AX[a2] = A x_cand; partials of ||dy||^2 and dy'(A dx).
template <typename T, typename AT> struct EpiProbe {
    static constexpr int kSlots = 2;
    Iter<T> I;
    const skk_state *st;
    skk_state *snap;    // or null
    double *part;       // region base
    int stride;         // part_capacity
    const double *dx_in;  // or null: the A^T product's ||dx||^2 partials to fold
    int dx_count;
    const T *y, *yc, *ax;
    T *axc;
    // Reads the buffer table from the kernel-argument copy `k`: selecting
    // among this object's own copy by a runtime index would place it in
    // scratch memory.
    template <class K> __device__ void begin(const K &k)
    {
        const int cur = st->cur;
        y = k.I.y(cur); yc = k.I.y(cur ^ 1); ax = k.I.ax(st->a0); axc = k.I.ax(st->a2);
    }
    static constexpr bool kPrefetch = sizeof(T) == 4;
    struct In {
        T yc, y, ax;
    };
    __device__ In load(int i) const { return In{yc[i], y[i], ax[i]}; }
    __device__ void finish(int i, AT s, const In &v, double *p) const
    {
        const T a = static_cast<T>(s);
        axc[i] = a;
        const double dy = static_cast<double>(v.yc) - static_cast<double>(v.y);
        p[0] += dy * dy;
        p[1] += dy * (static_cast<double>(a) - static_cast<double>(v.ax));
    }
    __device__ void row(int i, AT s, double *p) const { finish(i, s, load(i), p); }
    __device__ void end(double *p) const
    {
        double v[2] = {p[0], p[1]};
        store_partials<2>(v, part + (size_t)R_DY * stride, stride);
        // the state this launch ran with, for a decision fused into the
        // next launch (st itself is rewritten there)
        if (snap && blockIdx.x == 0 && threadIdx.x == 0) *snap = *st;
        // this workgroup's share of the A^T product's ||dx||^2 partials (g,
        // g + G, ... for G probe workgroups), summed by the first wave in a
        // fixed order: the decision then reduces one per probe workgroup
        if (dx_in && threadIdx.x < 64) {
            double a = 0.0;
            for (int k = blockIdx.x + static_cast<int>(threadIdx.x * gridDim.x); k < dx_count;
                 k += 64 * static_cast<int>(gridDim.x))
                a += dx_in[k];
            a = wave_sum(a);
            if (threadIdx.x == 0) part[(size_t)R_DXF * stride + blockIdx.x] = a;
        }
    }
};

// Accept/reject decision and next step size (PdlpKernels.cu:454-493) from
// the reduced probe sums, committed by rotating buffer indices: `out` is
// `in` after the decision.
// The step rule's factors after `updates` decisions: (1 - (k+1)^-0.3,
// 1 + (k+1)^-0.6) with k = updates + 1.  The decision kernels have one
// thread of a second wave compute them from the state loaded together with
// the probe partials, once those have arrived: the two pow calls then add
// no memory wait of their own.
struct StepFactors {
    double shrink, grow;
};

__device__ StepFactors step_factors(double updates)
{
    const double k = updates + 1.0;
    return StepFactors{1.0 - pow(k + 1.0, -0.3), 1.0 + pow(k + 1.0, -0.6)};
}

__device__ void decide(const skk_state &in, double dx2, double dq, double dy2, double dya, const StepFactors &f,
                       skk_state &out)
{
    out = in;
    // with a Hessian the local bound also covers dx'Q dx (>= 0 for convex Q)
    const double inter = fabs(dya) + 0.5 * fmax(dq, 0.0);
    const double omega = in.omega, eta = in.eta;
    const double movement = 0.5 * (omega * dx2 + dy2 / omega);
    const double eta_bar = (inter > 0.0 && movement > 0.0) ? movement / inter : in.eta_max;
    const bool accept = eta <= eta_bar;
    const double k = in.updates + 1.0;
    double next = fmin(f.shrink * eta_bar, f.grow * eta);
    if (!(next > 0.0) || !isfinite(next)) next = eta;
    next = fmin(fmax(next, in.eta_min), in.eta_max);
    out.eta = next;
    out.eta_bar = eta_bar;
    out.dx2 = dx2; out.dy2 = dy2; out.inter = inter;
    out.updates = k;
    out.accepted = accept ? 1 : 0;
    out.pending = accept ? 1 : 0;
    if (accept) {
        out.accept_count += 1.0;
        out.cur ^= 1;
        out.a0 = in.a2; out.a1 = in.a0; out.a2 = in.a1;   // x_prev <- x, x <- cand
    }
}

// The decision's four sums, reduced by one workgroup in a fixed order.
// kPer: partials per thread held in registers (counts above kPer * kBlock
// fall back to one slot at a time).
template <int kPer, class Mid>
__device__ void probe_sums(const double *part, int stride, int dx_blocks, int dq_blocks, int dy_blocks,
                           double &dx2, double &dq, double &dy2, double &dya, const Mid &mid,
                           int dx_region = R_DX)
{
    const double *const slots[4] = {part + (size_t)dx_region * stride, part + (size_t)R_DQ * stride,
                                    part + (size_t)R_DY * stride, part + (size_t)R_INTER * stride};
    const int counts[4] = {dx_blocks, dq_blocks, dy_blocks, dy_blocks};
    double out[4];
    combine_sums<4, kPer>(slots, counts, out, mid);
    dx2 = out[0];
    dq = dq_blocks > 0 ? out[1] : 0.0;
    dy2 = out[2];
    dya = out[3];
}

// One workgroup after the A x_cand launch: the decision, committed to st.
// dx_region: R_DX (the A^T product's partials) or R_DXF (folded by the
// probe launch, one per probe workgroup).
__global__ __launch_bounds__(kBlock) void k_probe_finalize(skk_state *st, const double *part, int stride,
                                                           int dx_blocks, int dq_blocks, int dy_blocks,
                                                           int dx_region)
{
    __shared__ StepFactors s_f;
    double dx2, dq, dy2, dya;
    // the state is loaded with the partials (one request per wave), so
    // neither the factors nor the decision wait for a load of their own;
    // the opaque zero makes these vector loads (scalar loads would be
    // waited for before the partials are requested)
    int zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
    const skk_state in = st[zero];
    auto factors = [&] {
        if (threadIdx.x == 64) s_f = step_factors(in.updates);   // visible after the tree's barriers
    };
    // as few partials per thread as the counts need (loads past a count are
    // issued too): the sums are the same for any kPer that covers them
    const int most = max(max(dx_blocks, dq_blocks), dy_blocks);
    if (most <= kBlock)
        probe_sums<1>(part, stride, dx_blocks, dq_blocks, dy_blocks, dx2, dq, dy2, dya, factors, dx_region);
    else if (most <= 4 * kBlock)
        probe_sums<4>(part, stride, dx_blocks, dq_blocks, dy_blocks, dx2, dq, dy2, dya, factors, dx_region);
    else
        probe_sums<SKK_SPMV_GRID / kBlock>(part, stride, dx_blocks, dq_blocks, dy_blocks, dx2, dq, dy2, dya,
                                           factors, dx_region);
    if (threadIdx.x != 0) return;
    skk_state next;
    decide(in, dx2, dq, dy2, dya, s_f, next);
    *st = next;
}

// The previous iteration's decision (from the snapshot its step probe took
// and the probe partials; every workgroup takes the same decision in the
// same order, workgroup 0 commits it to st) followed by this iteration's
// dual prox with the decided step and buffers.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dual_decide(Iter<T> I, Data<T> D, const skk_state *__restrict__ snap,
                                                       skk_state *st, const double *part, int stride,
                                                       int dx_blocks, int dq_blocks, int dy_blocks, int m)
{
    __shared__ skk_state s_next;
    __shared__ StepFactors s_f;
    double dx2, dq, dy2, dya;
    int zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
    const skk_state in = snap[zero];   // vector loads, in flight with the partials (k_probe_finalize)
    // This thread's first row with every buffer the decision can pick (y of
    // either slot, A x of all three), its bounds and running sum, requested
    // with the partials: the prox after the decision waits for no load.
    // (A row index past m reads row 0 -- every buffer holds at least one
    // element -- and is not stored.)
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = max(min(i0, m - 1), 0);
    const T py0 = I.Y[0][ic], py1 = I.Y[1][ic];
    const T pa0 = I.AX[0][ic], pa1 = I.AX[1][ic], pa2 = I.AX[2][ic];
    const T prl = D.rl[ic], pru = D.ru[ic];
    const T pyc = (D.yc ? D.yc : D.rl)[ic];
    const double psum = I.ysum.get(ic);
    // fused only while the products have at most SKK_FUSE_PARTIALS
    // workgroups together
    auto factors = [&] {
        if (threadIdx.x == 64) s_f = step_factors(in.updates);   // visible after the tree's barriers
    };
    if (max(max(dx_blocks, dq_blocks), dy_blocks) <= kBlock)
        probe_sums<1>(part, stride, dx_blocks, dq_blocks, dy_blocks, dx2, dq, dy2, dya, factors);
    else probe_sums<SKK_FUSE_PER>(part, stride, dx_blocks, dq_blocks, dy_blocks, dx2, dq, dy2, dya, factors);
    if (threadIdx.x == 0) {
        // decided in registers, then published (a decision written straight
        // into LDS would keep `in` in memory too, stored right after its loads)
        skk_state next;
        decide(in, dx2, dq, dy2, dya, s_f, next);
        s_next = next;
        if (blockIdx.x == 0) *st = next;
    }
    __syncthreads();
    const int cur = s_next.cur, pending = s_next.pending;
    const T sigma = static_cast<T>(s_next.eta * s_next.omega);
    const T *__restrict__ y = I.y(cur);
    const T *__restrict__ ax = I.ax(s_next.a0);
    const T *__restrict__ axp = I.ax(s_next.a1);
    T *__restrict__ yc = I.y(cur ^ 1);
    if (i0 < m) {
        // the same arithmetic as the loop below, on the prefetched values
        const T yi = cur ? py1 : py0;
        const int a0 = s_next.a0, a1 = s_next.a1;
        const T axv = a0 == 0 ? pa0 : (a0 == 1 ? pa1 : pa2);
        const T axpv = a1 == 0 ? pa0 : (a1 == 1 ? pa1 : pa2);
        const T v = yi + sigma * (T(2) * axv - axpv);
        yc[i0] = dual_prox_of(v, sigma, prl, pru, pyc, D.yc != nullptr);
        if (pending) {
            if (I.ysum.wide) static_cast<double *>(I.ysum.p)[i0] = psum + static_cast<double>(yi);
            else static_cast<T *>(I.ysum.p)[i0] = static_cast<T>(psum) + yi;
        }
    }
    const int stride_e = gridDim.x * blockDim.x;
    for (int i = i0 + stride_e; i < m; i += stride_e) {
        const T yi = y[i];
        const T v = yi + sigma * (T(2) * ax[i] - axp[i]);
        yc[i] = dual_prox(v, sigma, D, i);
        if (pending) I.ysum.add(i, yi);
    }
}

// Q x_cand epilogue: QX[cur^1] = Q x_cand; partial dx'(Q x_cand - Q x).
template <typename T, typename AT> struct EpiQProbe {
    static constexpr int kSlots = 1;
    Iter<T> I;
    const skk_state *st;
    double *part;
    const T *x, *xc, *qx;
    T *qxc;
    template <class K> __device__ void begin(const K &)
    {
        const int cur = st->cur;
        x = I.x(cur); xc = I.x(cur ^ 1); qx = I.qx(cur); qxc = I.qx(cur ^ 1);
    }
    static constexpr bool kPrefetch = sizeof(T) == 4;
    struct In {
        T xc, x, qx;
    };
    __device__ In load(int j) const { return In{xc[j], x[j], qx[j]}; }
    __device__ void finish(int j, AT s, const In &v, double *p) const
    {
        const T q = static_cast<T>(s);
        qxc[j] = q;
        p[0] += (static_cast<double>(v.xc) - static_cast<double>(v.x)) *
                (static_cast<double>(q) - static_cast<double>(v.qx));
    }
    __device__ void row(int j, AT s, double *p) const { finish(j, s, load(j), p); }
    __device__ void end(double *p) const
    {
        double v[1] = {p[0]};
        store_partials<1>(v, part, 0);
    }
};

// ---------------------------------------------------------- fixed PDHG step

// A xbar epilogue: y = prox(y + sigma A xbar), y_sum += y (PdlpKernels.cu:71-83).
template <typename T, typename AT> struct EpiDualFixed {
    static constexpr int kSlots = 0;
    Iter<T> I;
    Data<T> D;
    const skk_state *st;
    T *y;
    T sigma;
    template <class K> __device__ void begin(const K &)
    {
        y = I.y(st->cur);
        sigma = static_cast<T>(st->eta_fixed * st->omega);
    }
    __device__ void row(int i, AT s, double *) const
    {
        const T v = y[i] + sigma * static_cast<T>(s);
        const T yi = dual_prox(v, sigma, D, i);
        y[i] = yi;
        I.ysum.add(i, yi);
    }
    __device__ void end(double *) const {}
};

// A^T y epilogue: x+ = proj(x - tau (c + A^T y)), xbar = 2 x+ - x, x_sum += x+
// (PdlpKernels.cu:85-97).
template <typename T, typename AT> struct EpiPrimalFixed {
    static constexpr int kSlots = 0;
    Iter<T> I;
    Data<T> D;
    const skk_state *st;
    T *x;
    T tau;
    template <class K> __device__ void begin(const K &)
    {
        x = I.x(st->cur);
        tau = static_cast<T>(st->eta_fixed / st->omega);
    }
    __device__ void row(int j, AT s, double *) const
    {
        const T xj = x[j];
        const T xn = clampv(xj - tau * (D.c[j] + static_cast<T>(s)), D.l[j], D.u[j]);
        I.xbar[j] = T(2) * xn - xj;
        x[j] = xn;
        I.xsum.add(j, xn);
    }
    __device__ void end(double *) const {}
};

// ------------------------------------------------------------ evaluation

template <typename T>
__global__ void k_flush(Iter<T> I, const skk_state *__restrict__ st, int n, int m)
{
    if (!st->pending) return;
    const int cur = st->cur;
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) I.xsum.add(j, I.x(cur)[j]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) I.ysum.add(i, I.y(cur)[i]);
}

template <typename T>
__global__ void k_average(Iter<T> I, const skk_state *__restrict__ st, int n, int m, int adaptive,
                          double fixed_terms)
{
    double terms = fixed_terms;
    if (adaptive) {
        terms = st->accept_count - st->accepted_at_restart;
        if (terms < 1.0) terms = 1.0;
    }
    const T scale = static_cast<T>(1.0 / terms);
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) I.xavg[j] = static_cast<T>(scale * I.xsum.get(j));
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) I.yavg[i] = static_cast<T>(scale * I.ysum.get(i));
}

// The eight sums of PdlpKernels.cu:101-209 in FP64: ||Ax - proj Ax||^2,
// dual violation^2, c'x, dual objective, ||x||^2, ||y||^2, ||x - x_r||^2,
// ||y - y_r||^2.  Column terms use the FP64 data and FP64 products.
// `which` 0 evaluates X[cur], Y[cur]; 1 the running average.  With an
// anchor H the set holds corrections and the products are those of the
// corrections: the point evaluated is x = H.x + dx, y = H.y + dy, with
// A x = H.ax + A dx, c + A^T y + Q x = H.g + A^T dy + Q dx, Q x = H.qx + Q dx.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residuals_sel(Iter<T> I, int which,
                                                          const double *__restrict__ ax,
                                                          const double *__restrict__ aty,
                                                          Data<double> D, int n, int m,
                                                          double *part, int stride,
                                                          const skk_state *__restrict__ st,
                                                          const double *__restrict__ qx, Anchor H)
{
    const int cur = st->cur;
    const T *__restrict__ x = which ? I.xavg : I.x(cur);
    const T *__restrict__ y = which ? I.yavg : I.y(cur);
    const T *__restrict__ xr = I.xr;
    const T *__restrict__ yr = I.yr;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gs) {
        const double qd = qx ? qx[j] : 0.0;
        const double q = H.qx ? H.qx[j] + qd : qd;
        const double xd = x[j];
        const double xj = H.x ? H.x[j] + xd : xd, lo = D.l[j], hi = D.u[j];
        const double r = (H.x ? H.g[j] : D.c[j]) + aty[j] + qd;
        const double dx = xd - static_cast<double>(xr[j]);
        // with Q: primal c'x + x'Qx/2, dual terms of r = c + Qx + A'y minus x'Qx/2
        acc[2] += D.c[j] * xj + 0.5 * xj * q;
        acc[3] -= 0.5 * xj * q;
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
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gs) {
        const double a = H.x ? H.ax[i] + ax[i] : ax[i], rl = D.rl[i], ru = D.ru[i];
        const double pr = a - fmin(fmax(a, rl), ru);
        const double yd = y[i];
        const double yi = H.x ? H.y[i] + yd : yd, dy = yd - static_cast<double>(yr[i]);
        acc[0] += pr * pr;
        acc[5] += yi * yi;
        acc[7] += dy * dy;
        double support = 0.0;
        if (yi > 0.0) support = isfinite(ru) ? yi * ru : 0.0;
        else if (yi < 0.0) support = isfinite(rl) ? yi * rl : 0.0;
        acc[3] -= support;
    }
    store_partials<8>(acc, part, stride);
}

// ------------------------------------------- normalized duality gap (LP)
//
// The restart merit of Applegate et al. 2021 (s.3.1) for the evaluated
// point z = (x, y): G / r, G the largest gap L(x, yh) - L(xh, y) over the
// box and the omega-ball of radius r = ||z - z_r||_omega around z (z_r the
// last restart point).  The maximiser is one projected step of length t,
// xh = proj_[l,u](x - t g / omega), g = c + A^T y, yh = prox of the support
// function with step s = t omega, and the ball's distance grows with t:
// t is found by three rounds of 32 candidates (ratio 2 from the lower
// bound t0 = r / sqrt(|g|^2/omega + omega |Ax|^2), then 2^(1/31), then
// 2^(1/961)), each one pass over the vectors, and G is evaluated at the
// largest candidate inside the ball.  All in FP64.

constexpr int kNdgCand = 32;
enum { N_TBASE = 0, N_RATIO = 1, N_R2 = 2, N_R = 3 };

// The evaluated point in FP64 (anchor added) and the two norms of t0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ndg_prep(Iter<T> I, int which, const double *__restrict__ ax,
                                                     const double *__restrict__ aty, Data<double> D, Anchor H,
                                                     const skk_state *__restrict__ st, int n, int m, double *nx,
                                                     double *ng, double *ny, double *nax, double *part, int stride)
{
    const int cur = st->cur;
    const T *__restrict__ x = which ? I.xavg : I.x(cur);
    const T *__restrict__ y = which ? I.yavg : I.y(cur);
    double acc[2] = {0.0, 0.0};
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gs) {
        const double xd = x[j];
        const double g = (H.x ? H.g[j] : D.c[j]) + aty[j];
        nx[j] = H.x ? H.x[j] + xd : xd;
        ng[j] = g;
        acc[0] += g * g;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gs) {
        const double yd = y[i];
        const double a = H.x ? H.ax[i] + ax[i] : ax[i];
        ny[i] = H.x ? H.y[i] + yd : yd;
        nax[i] = a;
        acc[1] += a * a;
    }
    store_partials<2>(acc, part, stride);
}

// One workgroup: r from the evaluation's distance sums (slots 6, 7 of
// `sums`), t0, and the first search range.
__global__ __launch_bounds__(kBlock) void k_ndg_init(const double *part, int stride, int count,
                                                     const double *__restrict__ sums,
                                                     const skk_state *__restrict__ st, double *s)
{
    const double gx = combine(part, count, false), gy = combine(part + stride, count, false);
    if (threadIdx.x != 0) return;
    const double om = st->omega;
    const double r2 = om * fmax(0.0, sums[6]) + fmax(0.0, sums[7]) / om;
    const double gn = gx / om + om * gy;
    s[N_TBASE] = gn > 0.0 && r2 > 0.0 ? sqrt(r2 / gn) : 1.0;
    s[N_RATIO] = 2.0;
    s[N_R2] = r2;
    s[N_R] = sqrt(r2);
}

__device__ __forceinline__ double ndg_dual(double yv, double a, double s, double rl, double ru)
{
    // prox of the support function of [rl, ru] at v = y + s a, step s
    const double v = yv + s * a;
    return fmin(fmax(0.0, v - s * ru), v - s * rl);
}

__device__ __forceinline__ double ndg_support(double v, double lo, double hi)
{
    if (v > 0.0) return isfinite(hi) ? v * hi : 0.0;
    if (v < 0.0) return isfinite(lo) ? v * lo : 0.0;
    return 0.0;
}

// The squared omega-distance of the step for kNdgCand lengths at once.
__global__ __launch_bounds__(kBlock) void k_ndg_search(const double *__restrict__ nx, const double *__restrict__ ng,
                                                       const double *__restrict__ ny, const double *__restrict__ nax,
                                                       Data<double> D, const skk_state *__restrict__ st,
                                                       const double *__restrict__ s, int n, int m, double *part,
                                                       int stride)
{
    const double om = st->omega, t0 = s[N_TBASE], ratio = s[N_RATIO];
    double t[kNdgCand], acc[kNdgCand];
    t[0] = t0;
#pragma unroll
    for (int k = 1; k < kNdgCand; ++k) t[k] = t[k - 1] * ratio;
#pragma unroll
    for (int k = 0; k < kNdgCand; ++k) acc[k] = 0.0;
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gs) {
        const double xj = nx[j], gj = ng[j] / om, lo = D.l[j], hi = D.u[j];
#pragma unroll
        for (int k = 0; k < kNdgCand; ++k) {
            const double d = fmin(fmax(xj - t[k] * gj, lo), hi) - xj;
            acc[k] += om * d * d;
        }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gs) {
        const double yi = ny[i], a = nax[i], rl = D.rl[i], ru = D.ru[i];
#pragma unroll
        for (int k = 0; k < kNdgCand; ++k) {
            const double d = ndg_dual(yi, a, t[k] * om, rl, ru) - yi;
            acc[k] += d * d / om;
        }
    }
    store_partials<kNdgCand>(acc, part, stride);
}

// One workgroup: the largest candidate inside the ball starts the next,
// finer range (ratio^(1/31)); when every candidate is inside, the range
// moves on at the same ratio.
__global__ __launch_bounds__(kBlock) void k_ndg_pick(const double *part, int stride, int count, double *s)
{
    __shared__ double d2[kNdgCand];
    for (int k = 0; k < kNdgCand; ++k) {
        const double v = combine(part + (size_t)k * stride, count, false);
        if (threadIdx.x == 0) d2[k] = v;
    }
    if (threadIdx.x != 0) return;
    const double r2 = s[N_R2], t0 = s[N_TBASE], ratio = s[N_RATIO];
    int best = -1;
    for (int k = 0; k < kNdgCand; ++k) if (d2[k] <= r2) best = k;
    if (best < 0) {                                   // rounding at t0: step back
        s[N_TBASE] = t0 / ratio;
    } else if (best == kNdgCand - 1) {
        s[N_TBASE] = t0 * pow(ratio, (double)(kNdgCand - 1));
    } else {
        s[N_TBASE] = t0 * pow(ratio, (double)best);
        s[N_RATIO] = pow(ratio, 1.0 / (kNdgCand - 1));
    }
}

// The gap at the chosen step length, as differences (no objective-sized
// terms cancel): sum_j g_j (x_j - xh_j) + sum_i [(yh_i - y_i)(Ax)_i
// - p_i(yh_i) + p_i(y_i)].
__global__ __launch_bounds__(kBlock) void k_ndg_gap(const double *__restrict__ nx, const double *__restrict__ ng,
                                                    const double *__restrict__ ny, const double *__restrict__ nax,
                                                    Data<double> D, const skk_state *__restrict__ st,
                                                    const double *__restrict__ s, int n, int m, double *part,
                                                    int stride)
{
    const double om = st->omega, t = s[N_TBASE];
    double acc[1] = {0.0};
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gs) {
        const double xj = nx[j], g = ng[j];
        const double d = fmin(fmax(xj - t * g / om, D.l[j]), D.u[j]) - xj;
        acc[0] -= g * d;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gs) {
        const double yi = ny[i], a = nax[i], rl = D.rl[i], ru = D.ru[i];
        const double yh = ndg_dual(yi, a, t * om, rl, ru);
        acc[0] += (yh - yi) * a - ndg_support(yh, rl, ru) + ndg_support(yi, rl, ru);
    }
    store_partials<1>(acc, part, stride);
}

__global__ __launch_bounds__(kBlock) void k_ndg_value(const double *part, int count, const double *__restrict__ s,
                                                      double *out)
{
    const double G = combine(part, count, false);
    if (threadIdx.x != 0) return;
    const double r = s[N_R];
    *out = r > 0.0 && isfinite(G) ? G / r : (r > 0.0 ? INFINITY : 0.0);
}

template <typename T> int ndg_eval(const skk_pdhg &P, const Iter<T> &I, int which)
{
    int rc;
    const Data<double> D = data_of<double>(P);
    const int g = P.elem_grid, stride = P.part_capacity;
    hipLaunchKernelGGL(k_ndg_prep<T>, dim3(g), dim3(kBlock), 0, stream(), I, which, P.ax_d, P.aty_d, D,
                       anchor_of(P, std::is_same<T, float>::value), P.st, P.n, P.m, P.nx, P.ng, P.ny, P.nax,
                       P.ndg_part, stride);
    if ((rc = launch_status())) return rc;
    hipLaunchKernelGGL(k_ndg_init, dim3(1), dim3(kBlock), 0, stream(), P.ndg_part, stride, g,
                       P.eval_out + 8 * which, P.st, P.ndg_s);
    if ((rc = launch_status())) return rc;
    for (int round = 0; round < 3; ++round) {
        hipLaunchKernelGGL(k_ndg_search, dim3(g), dim3(kBlock), 0, stream(), P.nx, P.ng, P.ny, P.nax, D, P.st,
                           P.ndg_s, P.n, P.m, P.ndg_part, stride);
        if ((rc = launch_status())) return rc;
        hipLaunchKernelGGL(k_ndg_pick, dim3(1), dim3(kBlock), 0, stream(), P.ndg_part, stride, g, P.ndg_s);
        if ((rc = launch_status())) return rc;
    }
    hipLaunchKernelGGL(k_ndg_gap, dim3(g), dim3(kBlock), 0, stream(), P.nx, P.ng, P.ny, P.nax, D, P.st, P.ndg_s,
                       P.n, P.m, P.ndg_part, stride);
    if ((rc = launch_status())) return rc;
    hipLaunchKernelGGL(k_ndg_value, dim3(1), dim3(kBlock), 0, stream(), P.ndg_part, g, P.ndg_s,
                       P.eval_out + 16 + which);
    return launch_status();
}

// AX[a0] <- A x from the FP64 product of the evaluation.
template <typename T>
__global__ void k_to_t_sel(Iter<T> I, const double *__restrict__ src, int m,
                           const skk_state *__restrict__ st, const double *__restrict__ qsrc, int n)
{
    T *dst = I.ax(st->a0);
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) dst[i] = static_cast<T>(src[i]);
    if (qsrc) {
        T *q = I.qx(st->cur);
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) q[j] = static_cast<T>(qsrc[j]);
    }
}

// Restart (GpuPdlp.cpp:453-498 minus the primal-weight update, which the
// host applies through k_set_omega).
template <typename T>
__global__ void k_restart(Iter<T> I, skk_state *st, const double *__restrict__ ax_avg, int n, int m,
                          int use_avg, int adaptive, const double *__restrict__ qx_avg)
{
    const int cur = st->cur, a0 = st->a0, a1 = st->a1;
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        T x = I.x(cur)[j];
        if (use_avg) { x = I.xavg[j]; I.x(cur)[j] = x; if (qx_avg) I.qx(cur)[j] = static_cast<T>(qx_avg[j]); }
        I.xbar[j] = x;
        I.xsum.zero(j);
        I.xr[j] = x;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        T y = I.y(cur)[i];
        if (use_avg) { y = I.yavg[i]; I.y(cur)[i] = y; }
        I.ysum.zero(i);
        I.yr[i] = y;
        if (adaptive) {
            const T a = use_avg ? static_cast<T>(ax_avg[i]) : I.ax(a0)[i];
            I.ax(a0)[i] = a;
            I.ax(a1)[i] = a;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->accepted_at_restart = st->accept_count;
}

// FP32 -> FP64 copy of every iterate buffer (the state indices are shared).
// An anchored FP32 set holds corrections: the anchor is added to every
// point and product (the running sums are empty, a restart precedes this).
__global__ void k_promote(Iter<double> D, Iter<float> F, int n, int m, Anchor H)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const double h = H.x ? H.x[j] : 0.0, hq = H.qx ? H.qx[j] : 0.0;
        D.X[0][j] = h + F.X[0][j]; D.X[1][j] = h + F.X[1][j];
        static_cast<double *>(D.xsum.p)[j] = F.xsum.get(j);
        D.xr[j] = h + F.xr[j]; D.xbar[j] = h + F.xbar[j]; D.xavg[j] = h + F.xavg[j];
        if (D.QX[0]) { D.QX[0][j] = hq + F.QX[0][j]; D.QX[1][j] = hq + F.QX[1][j]; }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const double h = H.x ? H.y[i] : 0.0, ha = H.x ? H.ax[i] : 0.0;
        D.Y[0][i] = h + F.Y[0][i]; D.Y[1][i] = h + F.Y[1][i];
        D.AX[0][i] = ha + F.AX[0][i]; D.AX[1][i] = ha + F.AX[1][i]; D.AX[2][i] = ha + F.AX[2][i];
        static_cast<double *>(D.ysum.p)[i] = F.ysum.get(i); D.yr[i] = h + F.yr[i]; D.yavg[i] = h + F.yavg[i];
    }
}

// Anchored refinement, step 1: the anchor absorbs the current correction
// (a restart has just made it the restart point) and every correction
// buffer returns to zero.  The anchor's x is kept inside [l, u] (the
// exact iterate is; FP32 rounding of the shifted bounds may leave the sum
// an ulp outside).
__global__ void k_anchor_absorb(Iter<float> F, const skk_state *__restrict__ st, double *hx, double *hy,
                                const double *__restrict__ l, const double *__restrict__ u, int n, int m,
                                int absorb)
{
    const int cur = st->cur;
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        if (absorb) hx[j] = fmin(fmax(hx[j] + static_cast<double>(F.x(cur)[j]), l[j]), u[j]);
        F.X[0][j] = 0.0f; F.X[1][j] = 0.0f; F.xr[j] = 0.0f; F.xbar[j] = 0.0f; F.xavg[j] = 0.0f;
        F.xsum.zero(j);
        if (F.QX[0]) { F.QX[0][j] = 0.0f; F.QX[1][j] = 0.0f; }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        if (absorb) hy[i] += static_cast<double>(F.y(cur)[i]);
        F.Y[0][i] = 0.0f; F.Y[1][i] = 0.0f; F.yr[i] = 0.0f; F.yavg[i] = 0.0f;
        F.ysum.zero(i);
        F.AX[0][i] = 0.0f; F.AX[1][i] = 0.0f; F.AX[2][i] = 0.0f;
    }
}

// Anchored refinement, step 2: the correction problem's FP32 data from the
// FP64 products at the anchor (hax = A hx, hg = A^T hy on entry, hqx = Q hx).
// The column bounds are rounded inward, so a correction at its bound keeps
// hx + dx inside [l, u].
__global__ void k_anchor_data(const double *__restrict__ c, const double *__restrict__ l,
                              const double *__restrict__ u, const double *__restrict__ rl,
                              const double *__restrict__ ru, const double *__restrict__ hx,
                              const double *__restrict__ hy, const double *__restrict__ hax, double *hg,
                              const double *__restrict__ hqx, float *cf, float *lf, float *uf, float *rlf,
                              float *ruf, float *ycf, int n, int m)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const double g = c[j] + hg[j] + (hqx ? hqx[j] : 0.0);
        hg[j] = g;
        cf[j] = static_cast<float>(g);
        lf[j] = __double2float_ru(l[j] - hx[j]);
        uf[j] = __double2float_rd(u[j] - hx[j]);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        rlf[i] = static_cast<float>(rl[i] - hax[i]);
        ruf[i] = static_cast<float>(ru[i] - hax[i]);
        ycf[i] = static_cast<float>(hy[i]);
    }
}

// With an anchor the exported point is anchor + correction, x kept inside
// [l, u] (l, u: the FP64 column bounds).
template <typename T>
__global__ void k_export(Iter<T> I, const skk_state *__restrict__ st, int n, int m, int use_avg,
                         double *xo, double *yo, Anchor H, const double *__restrict__ l,
                         const double *__restrict__ u, skk_unpermute U)
{
    const int cur = st->cur;
    const int stride = gridDim.x * blockDim.x;
    // output position o of the original order holds slot U.col_slot[o]
    // (gathered), unscaled by the column factor: x = D_c x'
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += stride) {
        const int j = U.col_slot[o];
        const double v = static_cast<double>(use_avg ? I.xavg[j] : I.x(cur)[j]);
        xo[o] = (H.x ? fmin(fmax(H.x[j] + v, l[j]), u[j]) : v) * U.col_scale[o];
    }
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < m; o += stride) {
        const int i = U.row_slot[o];
        const double v = static_cast<double>(use_avg ? I.yavg[i] : I.y(cur)[i]);
        yo[o] = (H.x ? H.y[i] + v : v) * U.row_scale[o];
    }
}

// ------------------------------------------------------- power iteration

struct EpiNormSq {
    static constexpr int kSlots = 1;
    double *y, *part;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int r, double s, double *p) const { y[r] = s; p[0] += s * s; }
    __device__ void end(double *p) const
    {
        double v[1] = {p[0]};
        store_partials<1>(v, part, 0);
    }
};

__global__ void k_fill(double *v, double value, int n)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) v[j] = value;
}

__global__ void k_normalize_from(double *v, const double *src, const double *norm_sq, int n)
{
    const double nrm = sqrt(*norm_sq);
    const int stride = gridDim.x * blockDim.x;
    if (!(nrm > 0.0)) {
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) v[j] = src[j];
        return;
    }
    const double inv = 1.0 / nrm;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) v[j] = src[j] * inv;
}

__global__ void k_axpy(int n, double alpha, const double *x, double *y)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) y[j] = fma(alpha, x[j], y[j]);
}

// ---------------------------------------------------------- certificates

__device__ __forceinline__ bool sk_inf(double b) { return b >= 1.0e30; }
__device__ __forceinline__ bool sk_ninf(double b) { return b <= -1.0e30; }

// rc = c + Qx + A^T y epilogue for the certificate.
struct EpiRc {
    static constexpr int kSlots = 0;
    double *rc;
    const double *c, *qx;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int j, double s, double *) const { rc[j] = c[j] + (qx ? qx[j] : 0.0) + s; }
    __device__ void end(double *) const {}
};

// sk_verify (c/src/sk_core.c:394-510) residuals: sums pobj, dobj and
// x'Qx/2; maxima pinf, dinf, comp.
__global__ __launch_bounds__(kBlock) void k_verify(skk_verify_in V, int have_y)
{
    double s[3] = {0, 0, 0};   // c'x, dual objective terms, x'Qx
    double mx[3] = {0, 0, 0};  // pinf, dinf, comp
    bool bad_x = false;
    const int gs = gridDim.x * blockDim.x;
    const double bt = 1e-9;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < V.n; j += gs) {
        const double xj = V.x[j], lo = V.l[j], hi = V.u[j];
        s[0] += V.c[j] * xj;
        if (V.qx) s[2] += xj * V.qx[j];
        if (!isfinite(xj)) bad_x = true;
        if (!sk_ninf(lo) && lo - xj > mx[0]) mx[0] = lo - xj;
        if (!sk_inf(hi) && xj - hi > mx[0]) mx[0] = xj - hi;
        if (!have_y) continue;
        const double r = V.rc[j];
        const bool atlo = !sk_ninf(lo) && fabs(xj - lo) <= bt;
        const bool athi = !sk_inf(hi) && fabs(xj - hi) <= bt;
        if (atlo && athi) {
        } else if (atlo) { if (-r > mx[1]) mx[1] = -r; }
        else if (athi) { if (r > mx[1]) mx[1] = r; }
        else { if (fabs(r) > mx[1]) mx[1] = fabs(r); }
        if (r > 0.0) {
            if (sk_ninf(lo)) { if (r > mx[1]) mx[1] = r; } else s[1] += r * lo;
        } else if (r < 0.0) {
            if (sk_inf(hi)) { if (-r > mx[1]) mx[1] = -r; } else s[1] += r * hi;
        }
        if (r > 0.0 && !sk_ninf(lo)) { const double t = r * fabs(xj - lo); if (t > mx[2]) mx[2] = t; }
        else if (r < 0.0 && !sk_inf(hi)) { const double t = -r * fabs(hi - xj); if (t > mx[2]) mx[2] = t; }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V.m; i += gs) {
        const double a = V.act[i], rl = V.rl[i], ru = V.ru[i];
        if (!sk_ninf(rl) && rl - a > mx[0]) mx[0] = rl - a;
        if (!sk_inf(ru) && a - ru > mx[0]) mx[0] = a - ru;
        if (!have_y) continue;
        const double yv = V.y[i];
        const bool atlo = !sk_ninf(rl) && fabs(a - rl) <= bt;
        const bool athi = !sk_inf(ru) && fabs(a - ru) <= bt;
        if (atlo && athi) {
        } else if (atlo) { if (yv > mx[1]) mx[1] = yv; }
        else if (athi) { if (-yv > mx[1]) mx[1] = -yv; }
        else { if (fabs(yv) > mx[1]) mx[1] = fabs(yv); }
        if (yv < 0.0) {
            if (sk_ninf(rl)) { if (-yv > mx[1]) mx[1] = -yv; } else s[1] -= yv * rl;
            if (!sk_ninf(rl)) { const double t = -yv * fabs(a - rl); if (t > mx[2]) mx[2] = t; }
        } else if (yv > 0.0) {
            if (sk_inf(ru)) { if (yv > mx[1]) mx[1] = yv; } else s[1] -= yv * ru;
            if (!sk_inf(ru)) { const double t = yv * fabs(ru - a); if (t > mx[2]) mx[2] = t; }
        }
    }
    if (bad_x) mx[0] = INFINITY;
    store_partials<3>(s, V.part, V.grid);
    store_partials<3, true>(mx, V.part + 3 * (size_t)V.grid, V.grid);
}

// Relative-KKT sums in the original space (|b| >= 1e30 absent), with
// r = c + Qx + A^T y already in V.rc.
__global__ __launch_bounds__(kBlock) void k_relkkt(skk_verify_in V)
{
    double acc[5] = {0, 0, 0, 0, 0};
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < V.n; j += gs) {
        const double xj = V.x[j], r = V.rc[j];
        const double lo = sk_ninf(V.l[j]) ? -INFINITY : V.l[j];
        const double hi = sk_inf(V.u[j]) ? INFINITY : V.u[j];
        const double q = V.qx ? xj * V.qx[j] : 0.0;
        acc[2] += V.c[j] * xj;
        acc[4] += q;
        double viol = 0.0, term = 0.0;
        const bool lf = isfinite(lo), hf = isfinite(hi);
        if (lf && hf) term = fmin(r * lo, r * hi);
        else if (lf) { viol = fmax(0.0, -r); term = fmax(0.0, r) * lo; }
        else if (hf) { viol = fmax(0.0, r); term = fmin(0.0, r) * hi; }
        else viol = fabs(r);
        acc[1] += viol * viol;
        acc[3] += term;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V.m; i += gs) {
        const double a = V.act[i];
        const double rl = sk_ninf(V.rl[i]) ? -INFINITY : V.rl[i];
        const double ru = sk_inf(V.ru[i]) ? INFINITY : V.ru[i];
        const double pr = a - fmin(fmax(a, rl), ru);
        acc[0] += pr * pr;
        const double yi = V.y[i];
        double support = 0.0;
        if (yi > 0.0) support = isfinite(ru) ? yi * ru : 0.0;
        else if (yi < 0.0) support = isfinite(rl) ? yi * rl : 0.0;
        acc[3] -= support;
    }
    store_partials<5>(acc, V.part, V.grid);
}

// ------------------------------------------------ plain PDHG (legacy ABI)
//
// cuda/src/sankhya_cuda_lp.cu:26-52 semantics: per-row sigma and per-column
// tau (vectors or scalars), theta extrapolation, diagonal-Q implicit prox or
// an explicit Q x gradient.  FP64 throughout, like the reference.

__device__ __forceinline__ double project_box(double v, double lo, double hi)
{
    if (!isinf(lo) && v < lo) v = lo;
    if (!isinf(hi) && v > hi) v = hi;
    return v;
}

// A xbar epilogue: y <- trial - sigma proj(trial / sigma), in place.
struct EpiLegacyDual {
    static constexpr int kSlots = 0;
    double *y;
    const double *rl, *ru, *sigma_vec;
    double sigma;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int i, double s, double *) const
    {
        const double sg = sigma_vec ? sigma_vec[i] : sigma;
        const double trial = y[i] + sg * s;
        y[i] = trial - sg * project_box(trial / sg, rl[i], ru[i]);
    }
    __device__ void end(double *) const {}
};

// A^T y epilogue: x+ = proj((x - tau (c + A^T y + Qx)) / (1 + tau q)),
// xbar = x+ + theta (x+ - x), in place.
struct EpiLegacyPrimal {
    static constexpr int kSlots = 0;
    double *x, *xbar;
    const double *c, *l, *u, *tau_vec, *qg, *qd;
    double tau, theta;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int j, double s, double *) const
    {
        const double t = tau_vec ? tau_vec[j] : tau;
        const double curv = qd ? qd[j] : 0.0;
        const double qgr = qg ? qg[j] : 0.0;
        const double xj = x[j];
        const double trial = (xj - t * (c[j] + s + qgr)) / (1.0 + t * curv);
        const double next = project_box(trial, l[j], u[j]);
        x[j] = next;
        xbar[j] = next + theta * (next - xj);
    }
    __device__ void end(double *) const {}
};

// Checkpoint residuals of sankhya_cuda_lp.cu:258-367 on the device.
// Sums: objective, count of non-finite entries.  Maxima: step, scale,
// row violation, projected KKT, dual residual, complementarity.
struct LegacyCheck {
    int m, n, have_prev;
    const double *x, *y, *act, *grad, *qg, *qd, *c, *l, *u, *rl, *ru;
    double *prev;
    double *part; int stride; double *out;
};

__device__ __forceinline__ bool has_lo(double v) { return isfinite(v) && v > -5.0e29; }
__device__ __forceinline__ bool has_hi(double v) { return isfinite(v) && v < 5.0e29; }

__global__ __launch_bounds__(kBlock) void k_legacy_check(LegacyCheck K)
{
    double s[2] = {0.0, 0.0};
    double mx[6] = {0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    const int gs = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < K.n; j += gs) {
        const double xj = K.x[j], g = K.grad[j];
        const double qgr = K.qd ? K.qd[j] * xj : (K.qg ? K.qg[j] : 0.0);
        if (!isfinite(xj) || !isfinite(g) || (K.qg && !isfinite(K.qg[j]))) s[1] += 1.0;
        s[0] += K.c[j] * xj;
        if (K.qd) s[0] += 0.5 * K.qd[j] * xj * xj;
        else if (K.qg) s[0] += 0.5 * xj * K.qg[j];
        if (K.have_prev) mx[0] = fmax(mx[0], fabs(xj - K.prev[j]));
        K.prev[j] = xj;
        mx[1] = fmax(mx[1], fabs(xj));
        const double r = K.c[j] + g + qgr;
        const double lo = K.l[j], hi = K.u[j];
        mx[3] = fmax(mx[3], fabs(xj - project_box(xj - r, lo, hi)));
        const bool atlo = has_lo(lo) && fabs(xj - lo) <= 1e-9;
        const bool athi = has_hi(hi) && fabs(xj - hi) <= 1e-9;
        if (atlo && athi) {
        } else if (atlo) mx[4] = fmax(mx[4], -r);
        else if (athi) mx[4] = fmax(mx[4], r);
        else mx[4] = fmax(mx[4], fabs(r));
        if (r > 0.0 && has_lo(lo)) mx[5] = fmax(mx[5], r * fabs(xj - lo));
        else if (r < 0.0 && has_hi(hi)) mx[5] = fmax(mx[5], -r * fabs(hi - xj));
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < K.m; i += gs) {
        const double a = K.act[i], mu = K.y[i], rl = K.rl[i], ru = K.ru[i];
        if (!isfinite(a) || !isfinite(mu)) s[1] += 1.0;
        if (isfinite(rl)) mx[2] = fmax(mx[2], rl - a);
        if (isfinite(ru)) mx[2] = fmax(mx[2], a - ru);
        mx[3] = fmax(mx[3], fabs(a - project_box(a + mu, rl, ru)));
        const bool atlo = has_lo(rl) && fabs(a - rl) <= 1e-9;
        const bool athi = has_hi(ru) && fabs(a - ru) <= 1e-9;
        if (atlo && athi) {
        } else if (atlo) mx[4] = fmax(mx[4], mu);
        else if (athi) mx[4] = fmax(mx[4], -mu);
        else mx[4] = fmax(mx[4], fabs(mu));
        if (mu < 0.0 && has_lo(rl)) mx[5] = fmax(mx[5], -mu * fabs(a - rl));
        else if (mu > 0.0 && has_hi(ru)) mx[5] = fmax(mx[5], mu * fabs(ru - a));
    }
    store_partials<2>(s, K.part, K.stride);
    store_partials<6, true>(mx, K.part + 2 * (size_t)K.stride, K.stride);
}

__global__ void k_legacy_init(double *x, double *xbar, const double *l, const double *u, int n,
                              double *y, int m)
{
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const double v = project_box(0.0, l[j], u[j]);
        x[j] = v; xbar[j] = v;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) y[i] = 0.0;
}

int grid_for(int n, int cap)
{
    int g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return g < cap ? g : cap;
}


template <typename T> constexpr int prec_index() { return std::is_same<T, float>::value ? SKK_F32 : SKK_F64; }

// A x_cand with the probe epilogue (x_cand is X[cur ^ 1]).
// The probe launch folds the A^T product's ||dx||^2 partials when the
// one-workgroup decision follows it and the probe has fewer workgroups.
bool fold_dx(const skk_pdhg &P)
{
    return !P.fuse_decision && P.AT.grid > P.probe_grid;
}

template <typename T> void fold_into(const skk_pdhg &P, EpiProbe<T, T> &pr)
{
    if (fold_dx(P)) {
        pr.dx_in = P.part + (size_t)R_DX * P.part_capacity;
        pr.dx_count = P.AT.grid;
    }
}

template <typename T> int probe_spmv(const skk_pdhg &P, const Iter<T> &I, bool snapshot = false)
{
    EpiProbe<T, T> pr{};
    pr.I = I; pr.st = P.st; pr.part = P.part; pr.stride = P.part_capacity;
    pr.snap = snapshot ? P.snap : nullptr;
    fold_into(P, pr);
    const int kc = P.xcache_bytes / static_cast<int>(sizeof(T));
    if (kc <= 0 || P.A.ntiles == 0) return launch_spmv<T, T, T>(P.A, vals_of<T>(P.A), I.X[0], I.X[1], P.st, pr);
    // the last kc columns -- the longest, in slot order -- from LDS; the
    // copy starts and ends on 16-byte boundaries (columns past its end are
    // gathered from memory)
    constexpr int q = 16 / static_cast<int>(sizeof(T));
    const int c0a = ((P.n > kc ? P.n - kc : 0) + q - 1) / q * q;
    const int kcn = (P.n - c0a) / q * q;
    if (kcn <= 0) return launch_spmv<T, T, T>(P.A, vals_of<T>(P.A), I.X[0], I.X[1], P.st, pr);
    using E = EpiProbe<T, T>;
    const size_t lds = sizeof(T) * (size_t)kcn;
    if (P.A.stream)
        hipLaunchKernelGGL((k_spmv<T, T, T, E, true, true>), dim3(P.A.grid), dim3(kBlock), lds, stream(), P.A.rp,
                           P.A.ci, vals_of<T>(P.A), I.X[0], I.X[1], &P.st->cur, P.A.tiles, P.A.ntiles, pr, c0a, kcn);
    else
        hipLaunchKernelGGL((k_spmv<T, T, T, E, false, true>), dim3(P.A.grid), dim3(kBlock), lds, stream(), P.A.rp,
                           P.A.ci, vals_of<T>(P.A), I.X[0], I.X[1], &P.st->cur, P.A.tiles, P.A.ntiles, pr, c0a, kcn);
    return launch_status();
}

// Q x_cand and A x_cand in one launch of P.probe_grid workgroups (QP).
template <typename T> int qa_probe(const skk_pdhg &P, const Iter<T> &I, bool snapshot)
{
    EpiQProbe<T, T> eq{};
    eq.I = I; eq.st = P.st; eq.part = P.part + (size_t)R_DQ * P.part_capacity;
    EpiProbe<T, T> pr{};
    pr.I = I; pr.st = P.st; pr.part = P.part; pr.stride = P.part_capacity;
    pr.snap = snapshot ? P.snap : nullptr;
    fold_into(P, pr);
    const TileSet q{P.Q.rp, P.Q.ci, P.Q.tiles, P.Q.ntiles}, a{P.A.rp, P.A.ci, P.A.tiles, P.A.ntiles};
    const T *vq = vals_of<T>(P.Q), *va = vals_of<T>(P.A);
    using EQ = EpiQProbe<T, T>;
    using EA = EpiProbe<T, T>;
    const dim3 g(P.probe_grid), b(kBlock);
    if (P.Q.stream && P.A.stream)
        hipLaunchKernelGGL((k_spmv_pair<T, T, T, EQ, EA, true, true>), g, b, 0, stream(), q, vq, a, va, I.X[0], I.X[1], &P.st->cur, eq, pr);
    else if (P.Q.stream)
        hipLaunchKernelGGL((k_spmv_pair<T, T, T, EQ, EA, true, false>), g, b, 0, stream(), q, vq, a, va, I.X[0], I.X[1], &P.st->cur, eq, pr);
    else if (P.A.stream)
        hipLaunchKernelGGL((k_spmv_pair<T, T, T, EQ, EA, false, true>), g, b, 0, stream(), q, vq, a, va, I.X[0], I.X[1], &P.st->cur, eq, pr);
    else
        hipLaunchKernelGGL((k_spmv_pair<T, T, T, EQ, EA, false, false>), g, b, 0, stream(), q, vq, a, va, I.X[0], I.X[1], &P.st->cur, eq, pr);
    return launch_status();
}

// The step probe's products: A x_cand, with Q x_cand in the same launch for
// a QP.
template <typename T> int step_probe(const skk_pdhg &P, const Iter<T> &I, bool snapshot = false)
{
    return P.has_q ? qa_probe<T>(P, I, snapshot) : probe_spmv<T>(P, I, snapshot);
}

// The single-workgroup accept/reject decision.
int probe_decide(const skk_pdhg &P)
{
    const bool fold = fold_dx(P);
    hipLaunchKernelGGL(k_probe_finalize, dim3(1), dim3(kBlock), 0, stream(), P.st, P.part, P.part_capacity,
                       fold ? P.probe_grid : P.AT.grid, P.has_q ? P.probe_grid : 0, P.probe_grid,
                       fold ? static_cast<int>(R_DXF) : static_cast<int>(R_DX));
    return launch_status();
}

template <typename T> int probe(const skk_pdhg &P, const Iter<T> &I)
{
    const int rc = step_probe<T>(P, I);
    return rc ? rc : probe_decide(P);
}

template <typename T> int adaptive_iter(const skk_pdhg &P)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    const Data<T> D = data_of<T>(P);
    int rc;
    hipLaunchKernelGGL(k_dual_adaptive<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, D,
                       P.st, P.m);
    if ((rc = launch_status())) return rc;
    EpiPrimalAdaptive<T, T> ep{};
    ep.I = I; ep.D = D; ep.st = P.st;
    ep.part = P.part + (size_t)R_DX * P.part_capacity;
    ep.stride = P.part_capacity;
    // A^T y_cand: y_cand is Y[cur ^ 1], picked on the device
    if ((rc = launch_spmv<T, T, T>(P.AT, vals_of<T>(P.AT), I.Y[0], I.Y[1], P.st, ep))) return rc;
    return probe<T>(P, I);
}

// count iterations; with the fused decision: dual (first) or dual+decision,
// A^T + primal, [Q], A + probe (snapshot) per iteration, one closing decision.
template <typename T> int adaptive_run(const skk_pdhg &P, int count)
{
    if (!P.fuse_decision || !P.snap) {
        for (int k = 0; k < count; ++k) {
            const int rc = adaptive_iter<T>(P);
            if (rc) return rc;
        }
        return 0;
    }
    const Iter<T> I(P.it[prec_index<T>()]);
    const Data<T> D = data_of<T>(P);
    int rc;
    for (int k = 0; k < count; ++k) {
        if (k == 0) {
            hipLaunchKernelGGL(k_dual_adaptive<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, D, P.st, P.m);
        } else {
            hipLaunchKernelGGL(k_dual_decide<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, D, P.snap, P.st,
                               P.part, P.part_capacity, P.AT.grid, P.has_q ? P.probe_grid : 0, P.probe_grid, P.m);
        }
        if ((rc = launch_status())) return rc;
        EpiPrimalAdaptive<T, T> ep{};
        ep.I = I; ep.D = D; ep.st = P.st;
        ep.part = P.part + (size_t)R_DX * P.part_capacity;
        ep.stride = P.part_capacity;
        if ((rc = launch_spmv<T, T, T>(P.AT, vals_of<T>(P.AT), I.Y[0], I.Y[1], P.st, ep))) return rc;
        if ((rc = step_probe<T>(P, I, true))) return rc;
    }
    return count > 0 ? probe_decide(P) : 0;
}

template <typename T> int adaptive_part(const skk_pdhg &P, int part)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    const Data<T> D = data_of<T>(P);
    if (part == 0) {
        hipLaunchKernelGGL(k_dual_adaptive<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, D,
                           P.st, P.m);
        return launch_status();
    }
    if (part == 1) {
        EpiPrimalAdaptive<T, T> ep{};
        ep.I = I; ep.D = D; ep.st = P.st;
        ep.part = P.part + (size_t)R_DX * P.part_capacity;
        ep.stride = P.part_capacity;
        return launch_spmv<T, T, T>(P.AT, vals_of<T>(P.AT), I.Y[0], I.Y[1], P.st, ep);
    }
    if (part == 3) return probe_decide(P);
    return step_probe<T>(P, I);
}

template <typename T> int fixed_iter(const skk_pdhg &P)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    const Data<T> D = data_of<T>(P);
    int rc;
    EpiDualFixed<T, T> ed{};
    ed.I = I; ed.D = D; ed.st = P.st;
    if ((rc = launch_spmv<T, T, T>(P.A, vals_of<T>(P.A), I.xbar, I.xbar, nullptr, ed))) return rc;
    EpiPrimalFixed<T, T> ep{};
    ep.I = I; ep.D = D; ep.st = P.st;
    // the fixed path never rotates: cur stays 0
    return launch_spmv<T, T, T>(P.AT, vals_of<T>(P.AT), I.Y[0], I.Y[0], nullptr, ep);
}

template <typename T> int flush(const skk_pdhg &P)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    hipLaunchKernelGGL(k_flush<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, P.st, P.n, P.m);
    int rc = launch_status();
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear_pending, dim3(1), dim3(64), 0, stream(), P.st);
    return launch_status();
}

// Products in FP64 of the FP64 data with the current iterate (which 0) or
// the running average (which 1), then the eight residual sums.
template <typename T> int eval(const skk_pdhg &P, int which, int adaptive, double fixed_terms)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    int rc;
    const T *x, *y, *x1 = nullptr, *y1 = nullptr;
    const skk_state *sel = nullptr;
    if (which) {
        hipLaunchKernelGGL(k_average<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, P.st, P.n,
                           P.m, adaptive, fixed_terms);
        if ((rc = launch_status())) return rc;
        x = I.xavg; y = I.yavg;
        x1 = x; y1 = y;
    } else {
        // current iterate X[cur]: with sel the kernel reads X[cur ^ 1] of
        // (X[1], X[0]), i.e. X[cur]
        x = I.X[1]; x1 = I.X[0]; y = I.Y[1]; y1 = I.Y[0];
        sel = P.st;
    }
    EpiStore<double, double> sa{P.ax_d};
    if ((rc = launch_spmv<double, T, double>(P.A, P.A.vd, x, x1, sel, sa))) return rc;
    EpiStore<double, double> st{P.aty_d};
    if ((rc = launch_spmv<double, T, double>(P.AT, P.AT.vd, y, y1, sel, st))) return rc;
    if (P.has_q) {
        EpiStore<double, double> sq{P.qx_d};
        if ((rc = launch_spmv<double, T, double>(P.Q, P.Q.vd, x, x1, sel, sq))) return rc;
    }
    const Data<double> D = data_of<double>(P);
    double *rp = P.part + (size_t)R_EVAL * P.part_capacity;
    hipLaunchKernelGGL(k_residuals_sel<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, which,
                       P.ax_d, P.aty_d, D, P.n, P.m, rp, P.part_capacity, P.st,
                       P.has_q ? P.qx_d : nullptr, anchor_of(P, prec_index<T>() == SKK_F32));
    if ((rc = launch_status())) return rc;
    if ((rc = finalize(rp, P.part_capacity, P.elem_grid, 8, 0, P.eval_out + 8 * which))) return rc;
    if (P.ndg && (rc = ndg_eval<T>(P, I, which))) return rc;
    if (!which && adaptive) {
        // the live A x (and Q x) are refreshed from the FP64 products (GpuPdlp.cpp:367)
        hipLaunchKernelGGL(k_to_t_sel<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, P.ax_d,
                           P.m, P.st, P.has_q ? P.qx_d : nullptr, P.n);
        rc = launch_status();
    }
    return rc;
}

template <typename T> int restart(const skk_pdhg &P, int use_avg, int adaptive)
{
    const Iter<T> I(P.it[prec_index<T>()]);
    hipLaunchKernelGGL(k_restart<T>, dim3(P.elem_grid), dim3(kBlock), 0, stream(), I, P.st, P.ax_d,
                       P.n, P.m, use_avg, adaptive, P.has_q ? P.qx_d : nullptr);
    return launch_status();
}

// ---------------------------------------- relative KKT without allocation
//
// The relative KKT triple of an LP (GpuPdlp.cpp:171-176, 226-249) from two
// products whose epilogues reduce the residual terms directly, so only
// block partials need device memory: row i of A x gives the primal
// residual, the support term of y_i and the row-bound norms; row j of
// A^T y gives the reduced cost c_j + (A^T y)_j and the column terms.

struct EpiKktRow {
    static constexpr int kSlots = 4;
    const double *rl, *ru, *y;
    double *part;
    int stride;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int i, double a, double *p) const
    {
        const double lo = sk_ninf(rl[i]) ? -INFINITY : rl[i];
        const double hi = sk_inf(ru[i]) ? INFINITY : ru[i];
        const double pr = a - fmin(fmax(a, lo), hi);
        const double yi = y[i];
        double support = 0.0;
        if (yi > 0.0) support = isfinite(hi) ? yi * hi : 0.0;
        else if (yi < 0.0) support = isfinite(lo) ? yi * lo : 0.0;
        p[0] += pr * pr;
        p[1] -= support;
        if (isfinite(lo)) p[2] += lo * lo;
        if (isfinite(hi)) p[3] += hi * hi;
    }
    __device__ void end(double *p) const
    {
        double v[4] = {p[0], p[1], p[2], p[3]};
        store_partials<4>(v, part, stride);
    }
};

struct EpiKktCol {
    static constexpr int kSlots = 4;
    const double *c, *l, *u, *x;
    double *part;
    int stride;
    template <class K> __device__ void begin(const K &) {}
    __device__ void row(int j, double s, double *p) const
    {
        const double r = c[j] + s, xj = x[j];
        const double lo = sk_ninf(l[j]) ? -INFINITY : l[j];
        const double hi = sk_inf(u[j]) ? INFINITY : u[j];
        double viol = 0.0, term = 0.0;
        const bool lf = isfinite(lo), hf = isfinite(hi);
        if (lf && hf) term = fmin(r * lo, r * hi);
        else if (lf) { viol = fmax(0.0, -r); term = fmax(0.0, r) * lo; }
        else if (hf) { viol = fmax(0.0, r); term = fmin(0.0, r) * hi; }
        else viol = fabs(r);
        p[0] += viol * viol;
        p[1] += c[j] * xj;
        p[2] += term;
        p[3] += c[j] * c[j];
    }
    __device__ void end(double *p) const
    {
        double v[4] = {p[0], p[1], p[2], p[3]};
        store_partials<4>(v, part, stride);
    }
};

// One workgroup: the two products' partials in a fixed order, then
// out[0] primal, [1] dual, [2] gap, [3] primal objective, [4] dual objective.
__global__ __launch_bounds__(kBlock) void k_kkt_finalize(const double *rows, int nrow_blocks, const double *cols,
                                                         int ncol_blocks, int stride, double *out)
{
    double t[8];
    for (int s = 0; s < 4; ++s) t[s] = combine(rows + (size_t)s * stride, nrow_blocks, false);
    for (int s = 0; s < 4; ++s) t[4 + s] = combine(cols + (size_t)s * stride, ncol_blocks, false);
    if (threadIdx.x != 0) return;
    const double bound_norm = sqrt(fmax(0.0, t[2] + t[3]));
    const double cost_norm = sqrt(fmax(0.0, t[7]));
    double primal = sqrt(fmax(0.0, t[0])) / (1.0 + bound_norm);
    double dual = sqrt(fmax(0.0, t[4])) / (1.0 + cost_norm);
    const double pobj = t[5], dobj = t[6] + t[1];
    double gap = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    if (!isfinite(primal) || !isfinite(dual) || !isfinite(gap)) primal = dual = gap = INFINITY;
    out[0] = primal; out[1] = dual; out[2] = gap; out[3] = pobj; out[4] = dobj;
}

}  // namespace

// ======================================================== extern "C" layer


// --------------------------------------------------- gather-rate probe

// Index stream for the gather probe: a SplitMix64 hash of the position,
// reduced to [0, table) -- uniform, reproducible, no host transfer.
__global__ void k_probe_fill(int *idx, float *val, int count, int table, float *tab)
{
    const int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += stride) {
        uint64_t z = (uint64_t)p * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        idx[p] = (int)(z % (uint64_t)table);
        val[p] = 1.0f;
        if (p < table) tab[p] = 0.5f;
    }
}

// The access pattern of a CSR product without its row bookkeeping: lanes
// stride over the (value, index) stream with coalesced loads, eight pairs
// and their eight gathers table[index] in flight per lane, over a grid of
// at most SKK_SPMV_GRID workgroups (the products' own grid cap).
__global__ __launch_bounds__(kBlock) void k_probe_gather(const int *__restrict__ idx, const float *__restrict__ val,
                                                         const float *__restrict__ tab, float *out, int count)
{
    const int stride = gridDim.x * kBlock;
    int k = blockIdx.x * kBlock + threadIdx.x;
    float s = 0.0f;
    for (; k + 7 * stride < count; k += 8 * stride) {
        int j[8];
        float v[8], g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) { j[u] = idx[k + u * stride]; v[u] = val[k + u * stride]; }
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = tab[j[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u] * g[u];
    }
    for (; k < count; k += stride) s += val[k] * tab[idx[k]];
    out[blockIdx.x * kBlock + threadIdx.x] = s;
}

// Device scratch of skk_relkkt_noalloc: the block partials of its two
// products, 4 slots of at most SKK_SPMV_GRID workgroups each (256 KB).
// Every GPU holds its own copy until the process ends; the kernels take
// its device address, looked up once per device.
constexpr int kNoAllocSlots = 4;
constexpr int kNoAllocRegion = kNoAllocSlots * SKK_SPMV_GRID;
static __device__ __align__(256) double scratch[2 * kNoAllocRegion];
static double *g_scratch[kMaxDevices] = {};

extern "C" {

int skk_device_count(int *count)
{
    *count = 0;
    return check(hipGetDeviceCount(count));
}

int skk_malloc(void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 8;
    return check(hipMalloc(ptr, bytes));
}

void skk_free(void *ptr)
{
    if (ptr) (void)hipFree(ptr);
}

int skk_h2d(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    int rc = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    return rc ? rc : check(hipStreamSynchronize(stream()));
}

int skk_d2h(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    int rc = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    return rc ? rc : check(hipStreamSynchronize(stream()));
}

int skk_d2h_async(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
}

int skk_d2d(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
}

int skk_host_alloc(void **ptr, size_t bytes)
{
    *ptr = nullptr;
    return check(hipHostMalloc(ptr, bytes ? bytes : 8, hipHostMallocDefault));
}

void skk_host_free(void *ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

int skk_sync(void) { return check(hipStreamSynchronize(stream())); }

int skk_last_launch_error(void) { return g_last; }

void skk_reset_error(void)
{
    (void)hipGetLastError();   // drop a runtime error left by an earlier, already reported call
    g_last = 0;
}

const char *skk_error_text(int code) { return hipGetErrorString(static_cast<hipError_t>(code)); }

void *skk_stream(void) { return static_cast<void *>(stream()); }

int skk_graph_begin(void)
{
    return check(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
}

int skk_graph_end(void **exec)
{
    hipGraph_t g = nullptr;
    *exec = nullptr;
    int rc = check(hipStreamEndCapture(stream(), &g));
    if (rc) return rc;
    hipGraphExec_t e = nullptr;
    rc = check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    *exec = static_cast<void *>(e);
    return 0;
}

int skk_graph_launch(void *exec)
{
    return check(hipGraphLaunch(static_cast<hipGraphExec_t>(exec), stream()));
}

void skk_graph_destroy(void *exec)
{
    if (exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec));
}

int skk_mark(int slot)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= kMaxDevices || slot < 0 || slot >= 8) return -1;
    if (!g_marks[dev][slot] && check(hipEventCreate(&g_marks[dev][slot]))) return g_last;
    return check(hipEventRecord(g_marks[dev][slot], stream()));
}

double skk_elapsed_ms(int a, int b)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    float ms = 0.0f;
    if (dev < 0 || dev >= kMaxDevices || !g_marks[dev][a] || !g_marks[dev][b]) return -1.0;
    if (hipEventSynchronize(g_marks[dev][b]) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, g_marks[dev][a], g_marks[dev][b]) != hipSuccess) return -1.0;
    return static_cast<double>(ms);
}

int skk_spmv_f64(const skk_csr *A, const double *x, double *y)
{
    EpiStore<double, double> e{y};
    return launch_spmv<double, double, double>(*A, A->vd, x, x, nullptr, e);
}

int skk_spmv_f32(const skk_csr *A, const float *x, float *y)
{
    if (!A->vf) return -1;
    EpiStore<float, float> e{y};
    return launch_spmv<float, float, float>(*A, A->vf, x, x, nullptr, e);
}

int skk_axpy_f64(int n, double alpha, const double *x, double *y)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 2048)), dim3(kBlock), 0, stream(), n, alpha, x, y);
    return launch_status();
}

int skk_pdhg_init(const skk_pdhg *P, int prec, double eta0, double omega, double eta_min,
                  double eta_max, int warm_x, int warm_y)
{
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, stream(), P->st, eta0, omega, eta_min,
                       eta_max);
    int rc = launch_status();
    if (rc) return rc;
    if (prec == SKK_F32) {
        Iter<float> I(P->it[SKK_F32]);
        hipLaunchKernelGGL(k_init_vectors<float>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I,
                           P->n, P->m, warm_x, warm_y);
    } else {
        Iter<double> I(P->it[SKK_F64]);
        hipLaunchKernelGGL(k_init_vectors<double>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I,
                           P->n, P->m, warm_x, warm_y);
    }
    if ((rc = launch_status())) return rc;
    if (warm_x) {
        // A x0 into AX[0] and AX[1] (extrapolation starts as the identity)
        if (prec == SKK_F32) {
            Iter<float> I(P->it[SKK_F32]);
            EpiStore<float, float> e{I.AX[0]};
            if ((rc = launch_spmv<float, float, float>(P->A, P->A.vf, I.X[0], I.X[0], nullptr, e))) return rc;
            hipLaunchKernelGGL(k_copy<float>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I.AX[1], I.AX[0], P->m);
            if (P->has_q) {
                EpiStore<float, float> q{I.QX[0]};
                if ((rc = launch_spmv<float, float, float>(P->Q, P->Q.vf, I.X[0], I.X[0], nullptr, q))) return rc;
                hipLaunchKernelGGL(k_copy<float>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I.QX[1], I.QX[0], P->n);
            }
        } else {
            Iter<double> I(P->it[SKK_F64]);
            EpiStore<double, double> e{I.AX[0]};
            if ((rc = launch_spmv<double, double, double>(P->A, P->A.vd, I.X[0], I.X[0], nullptr, e))) return rc;
            hipLaunchKernelGGL(k_copy<double>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I.AX[1], I.AX[0], P->m);
            if (P->has_q) {
                EpiStore<double, double> q{I.QX[0]};
                if ((rc = launch_spmv<double, double, double>(P->Q, P->Q.vd, I.X[0], I.X[0], nullptr, q))) return rc;
                hipLaunchKernelGGL(k_copy<double>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I.QX[1], I.QX[0], P->n);
            }
        }
        rc = launch_status();
    }
    return rc;
}

int skk_pdhg_adaptive_iter(const skk_pdhg *P, int prec)
{
    return prec == SKK_F32 ? adaptive_iter<float>(*P) : adaptive_iter<double>(*P);
}

int skk_pdhg_adaptive_run(const skk_pdhg *P, int prec, int count)
{
    return prec == SKK_F32 ? adaptive_run<float>(*P, count) : adaptive_run<double>(*P, count);
}

int skk_pdhg_adaptive_part(const skk_pdhg *P, int prec, int part)
{
    return prec == SKK_F32 ? adaptive_part<float>(*P, part) : adaptive_part<double>(*P, part);
}

static hipEvent_t *g_timeline[kMaxDevices] = {};
static int g_timeline_n[kMaxDevices] = {};

int skk_timeline_reserve(int count)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= kMaxDevices || count < 0) return -1;
    if (count <= g_timeline_n[dev]) return 0;
    hipEvent_t *ev = static_cast<hipEvent_t *>(realloc(g_timeline[dev], sizeof(hipEvent_t) * count));
    if (!ev) return -1;
    g_timeline[dev] = ev;
    for (int k = g_timeline_n[dev]; k < count; ++k) {
        if (check(hipEventCreate(&ev[k]))) { g_timeline_n[dev] = k; return g_last; }
    }
    g_timeline_n[dev] = count;
    return 0;
}

int skk_timeline_mark(int k)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= kMaxDevices || k < 0 || k >= g_timeline_n[dev]) return -1;
    return check(hipEventRecord(g_timeline[dev][k], stream()));
}

double skk_timeline_ms(int a, int b)
{
    int dev = 0;
    float ms = 0.0f;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= kMaxDevices || a < 0 || b < 0 || a >= g_timeline_n[dev] || b >= g_timeline_n[dev])
        return -1.0;
    if (hipEventSynchronize(g_timeline[dev][b]) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, g_timeline[dev][a], g_timeline[dev][b]) != hipSuccess) return -1.0;
    return static_cast<double>(ms);
}

int skk_pdhg_fixed_iter(const skk_pdhg *P, int prec)
{
    return prec == SKK_F32 ? fixed_iter<float>(*P) : fixed_iter<double>(*P);
}

int skk_pdhg_flush(const skk_pdhg *P, int prec)
{
    return prec == SKK_F32 ? flush<float>(*P) : flush<double>(*P);
}

int skk_pdhg_eval(const skk_pdhg *P, int prec, int which, int adaptive, double fixed_terms)
{
    return prec == SKK_F32 ? eval<float>(*P, which, adaptive, fixed_terms)
                           : eval<double>(*P, which, adaptive, fixed_terms);
}

int skk_pdhg_restart(const skk_pdhg *P, int prec, int use_average, int adaptive)
{
    return prec == SKK_F32 ? restart<float>(*P, use_average, adaptive)
                           : restart<double>(*P, use_average, adaptive);
}

int skk_pdhg_set_omega(const skk_pdhg *P, double omega)
{
    hipLaunchKernelGGL(k_set_omega, dim3(1), dim3(64), 0, stream(), P->st, omega);
    return launch_status();
}

int skk_pdhg_set_eta(const skk_pdhg *P, double eta)
{
    hipLaunchKernelGGL(k_set_eta, dim3(1), dim3(64), 0, stream(), P->st, eta);
    return launch_status();
}

int skk_pdhg_promote(const skk_pdhg *P)
{
    Iter<double> D(P->it[SKK_F64]);
    Iter<float> F(P->it[SKK_F32]);
    hipLaunchKernelGGL(k_promote, dim3(P->elem_grid), dim3(kBlock), 0, stream(), D, F, P->n, P->m,
                       anchor_of(*P, true));
    return launch_status();
}

int skk_pdhg_reanchor(const skk_pdhg *P, int absorb)
{
    if (!P->anchored) return -1;
    int rc;
    Iter<float> F(P->it[SKK_F32]);
    hipLaunchKernelGGL(k_anchor_absorb, dim3(P->elem_grid), dim3(kBlock), 0, stream(), F, P->st, P->hx, P->hy,
                       P->ld, P->ud, P->n, P->m, absorb);
    if ((rc = launch_status())) return rc;
    // FP64 products at the anchor; an empty operator leaves its product 0
    hipLaunchKernelGGL(k_fill, dim3(grid_for(P->m, P->elem_grid)), dim3(kBlock), 0, stream(), P->hax, 0.0, P->m);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(P->n, P->elem_grid)), dim3(kBlock), 0, stream(), P->hg, 0.0, P->n);
    if ((rc = launch_status())) return rc;
    EpiStore<double, double> ea{P->hax};
    if ((rc = launch_spmv<double, double, double>(P->A, P->A.vd, P->hx, P->hx, nullptr, ea))) return rc;
    EpiStore<double, double> eg{P->hg};
    if ((rc = launch_spmv<double, double, double>(P->AT, P->AT.vd, P->hy, P->hy, nullptr, eg))) return rc;
    if (P->has_q) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(P->n, P->elem_grid)), dim3(kBlock), 0, stream(), P->hqx, 0.0, P->n);
        if ((rc = launch_status())) return rc;
        EpiStore<double, double> eq{P->hqx};
        if ((rc = launch_spmv<double, double, double>(P->Q, P->Q.vd, P->hx, P->hx, nullptr, eq))) return rc;
    }
    hipLaunchKernelGGL(k_anchor_data, dim3(P->elem_grid), dim3(kBlock), 0, stream(), P->cd, P->ld, P->ud, P->rld,
                       P->rud, P->hx, P->hy, P->hax, P->hg, P->has_q ? P->hqx : nullptr,
                       const_cast<float *>(P->cf), const_cast<float *>(P->lf), const_cast<float *>(P->uf),
                       const_cast<float *>(P->rlf), const_cast<float *>(P->ruf), P->ycf, P->n, P->m);
    return launch_status();
}

int skk_pdhg_export(const skk_pdhg *P, int prec, int use_average, const skk_unpermute *U, double *x_out,
                    double *y_out)
{
    if (prec == SKK_F32) {
        Iter<float> I(P->it[SKK_F32]);
        hipLaunchKernelGGL(k_export<float>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I, P->st,
                           P->n, P->m, use_average, x_out, y_out, anchor_of(*P, true), P->ld, P->ud, *U);
    } else {
        Iter<double> I(P->it[SKK_F64]);
        hipLaunchKernelGGL(k_export<double>, dim3(P->elem_grid), dim3(kBlock), 0, stream(), I, P->st,
                           P->n, P->m, use_average, x_out, y_out, anchor_of(*P, false), P->ld, P->ud, *U);
    }
    return launch_status();
}

int skk_fill_f64(double *v, double value, int n)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 2048)), dim3(kBlock), 0, stream(), v, value, n);
    return launch_status();
}

int skk_power_norm(const skk_pdhg *P, double *v, double *w, double *v2, int iters, double *out)
{
    int rc;
    const int n = P->n;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, P->elem_grid)), dim3(kBlock), 0, stream(), v,
                       1.0 * (1.0 / sqrt(static_cast<double>(n > 0 ? n : 1))), n);
    if ((rc = launch_status())) return rc;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream(), out, 0.0, 1);
    if ((rc = launch_status())) return rc;
    for (int k = 0; k < iters; ++k) {
        EpiStore<double, double> e{w};
        if ((rc = launch_spmv<double, double, double>(P->A, P->A.vd, v, v, nullptr, e))) return rc;
        EpiNormSq ns{};
        ns.y = v2; ns.part = P->part + (size_t)R_POWER * P->part_capacity;
        if ((rc = launch_spmv<double, double, double>(P->AT, P->AT.vd, w, w, nullptr, ns))) return rc;
        if ((rc = finalize(ns.part, P->part_capacity, P->AT.grid, 1, 0, out))) return rc;
        hipLaunchKernelGGL(k_normalize_from, dim3(grid_for(n, P->elem_grid)), dim3(kBlock), 0,
                           stream(), v, v2, out, n);
        if ((rc = launch_status())) return rc;
    }
    return 0;
}

int skk_verify(const skk_verify_in *V)
{
    int rc;
    EpiStore<double, double> ea{V->act};
    if ((rc = launch_spmv<double, double, double>(V->A, V->A.vd, V->x, V->x, nullptr, ea))) return rc;
    if (V->Q) {
        EpiStore<double, double> eq{V->qx};
        if ((rc = launch_spmv<double, double, double>(*V->Q, V->Q->vd, V->x, V->x, nullptr, eq))) return rc;
    }
    if (V->y) {
        EpiRc er{V->rc, V->c, V->Q ? V->qx : nullptr};
        if (V->AT.rows > 0) {
            if ((rc = launch_spmv<double, double, double>(V->AT, V->AT.vd, V->y, V->y, nullptr, er))) return rc;
        }
    }
    skk_verify_in v = *V;
    if (!V->Q) v.qx = nullptr;
    hipLaunchKernelGGL(k_verify, dim3(V->grid), dim3(kBlock), 0, stream(), v, V->y ? 1 : 0);
    if ((rc = launch_status())) return rc;
    return finalize(V->part, V->grid, V->grid, 3, 3, V->out);
}

int skk_relative_kkt(const skk_verify_in *V)
{
    skk_verify_in v = *V;
    if (!V->Q) v.qx = nullptr;
    hipLaunchKernelGGL(k_relkkt, dim3(V->grid), dim3(kBlock), 0, stream(), v);
    int rc = launch_status();
    if (rc) return rc;
    return finalize(V->part, V->grid, V->grid, 5, 0, V->out);
}


int skk_legacy_init(double *x, double *xbar, const double *l, const double *u, int n, double *y, int m)
{
    const int g = grid_for(n > m ? n : m, 2048);
    hipLaunchKernelGGL(k_legacy_init, dim3(g), dim3(kBlock), 0, stream(), x, xbar, l, u, n, y, m);
    return launch_status();
}

int skk_legacy_iter(const skk_legacy *L)
{
    int rc;
    EpiLegacyDual ed{L->y, L->rl, L->ru, L->sigma_vec, L->sigma};
    if ((rc = launch_spmv<double, double, double>(*L->A, L->A->vd, L->xbar, L->xbar, nullptr, ed))) return rc;
    if (L->Q) {
        EpiStore<double, double> eq{L->qg};
        if ((rc = launch_spmv<double, double, double>(*L->Q, L->Q->vd, L->x, L->x, nullptr, eq))) return rc;
    }
    EpiLegacyPrimal ep{L->x, L->xbar, L->c, L->l, L->u, L->tau_vec, L->Q ? L->qg : nullptr, L->qd,
                       L->tau, L->theta};
    return launch_spmv<double, double, double>(*L->AT, L->AT->vd, L->y, L->y, nullptr, ep);
}

int skk_legacy_check(const skk_legacy *L, int have_prev)
{
    int rc;
    EpiStore<double, double> ea{L->act};
    if ((rc = launch_spmv<double, double, double>(*L->A, L->A->vd, L->x, L->x, nullptr, ea))) return rc;
    EpiStore<double, double> eg{L->grad};
    if ((rc = launch_spmv<double, double, double>(*L->AT, L->AT->vd, L->y, L->y, nullptr, eg))) return rc;
    if (L->Q) {
        EpiStore<double, double> eq{L->qg};
        if ((rc = launch_spmv<double, double, double>(*L->Q, L->Q->vd, L->x, L->x, nullptr, eq))) return rc;
    }
    LegacyCheck K{L->m, L->n, have_prev, L->x, L->y, L->act, L->grad, L->Q ? L->qg : nullptr, L->qd,
                  L->c, L->l, L->u, L->rl, L->ru, L->prev, L->part, L->grid, L->out};
    hipLaunchKernelGGL(k_legacy_check, dim3(L->grid), dim3(kBlock), 0, stream(), K);
    if ((rc = launch_status())) return rc;
    return finalize(L->part, L->grid, L->grid, 2, 6, L->out);
}

int skk_relkkt_noalloc(const skk_csr *A, const skk_csr *AT, const double *c, const double *rl, const double *ru,
                       const double *l, const double *u, const double *x, const double *y, double *out)
{
    int dev = 0;
    if (check(hipGetDevice(&dev))) return g_last;
    if (dev < 0 || dev >= kMaxDevices) return -1;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return g_last;
        g_scratch[dev] = static_cast<double *>(p);
    }
    if (A->grid > SKK_SPMV_GRID || AT->grid > SKK_SPMV_GRID) return -1;
    double *rows = g_scratch[dev], *cols = g_scratch[dev] + kNoAllocRegion;
    int rc;
    EpiKktRow er{rl, ru, y, rows, SKK_SPMV_GRID};
    if ((rc = launch_spmv<double, double, double>(*A, A->vd, x, x, nullptr, er))) return rc;
    EpiKktCol ec{c, l, u, x, cols, SKK_SPMV_GRID};
    if ((rc = launch_spmv<double, double, double>(*AT, AT->vd, y, y, nullptr, ec))) return rc;
    hipLaunchKernelGGL(k_kkt_finalize, dim3(1), dim3(kBlock), 0, stream(), rows, A->ntiles ? A->grid : 0, cols,
                       AT->ntiles ? AT->grid : 0, SKK_SPMV_GRID, out);
    return launch_status();
}

/* Random-gather issue rate of this device: `count` 4-byte gathers from a
 * table of `table` floats (uniform random indices), mean of `reps` timed
 * launches after one warm-up.  Returns gathers per second in *rate. */
int skk_gather_probe(int table, int count, int reps, double *rate)
{
    if (table <= 0 || count <= 0 || count > (1 << 30) || reps <= 0 || !rate) return -1;
    int *idx = nullptr;
    float *val = nullptr, *tab = nullptr, *out = nullptr;
    const int grid = (int)(((long long)count + kBlock - 1) / kBlock < SKK_SPMV_GRID
                           ? ((long long)count + kBlock - 1) / kBlock : SKK_SPMV_GRID);
    int rc = skk_malloc((void **)&idx, sizeof(int) * (size_t)count);
    if (!rc) rc = skk_malloc((void **)&val, sizeof(float) * (size_t)count);
    if (!rc) rc = skk_malloc((void **)&tab, sizeof(float) * (size_t)table);
    if (!rc) rc = skk_malloc((void **)&out, sizeof(float) * (size_t)grid * kBlock);
    if (!rc) {
        hipLaunchKernelGGL(k_probe_fill, dim3(1024), dim3(kBlock), 0, stream(), idx, val, count, table, tab);
        hipLaunchKernelGGL(k_probe_gather, dim3(grid), dim3(kBlock), 0, stream(), idx, val, tab, out, count);
        rc = launch_status();
    }
    if (!rc) rc = skk_timeline_reserve(2);
    if (!rc) rc = skk_timeline_mark(0);
    for (int r = 0; r < reps && !rc; ++r) {
        hipLaunchKernelGGL(k_probe_gather, dim3(grid), dim3(kBlock), 0, stream(), idx, val, tab, out, count);
        rc = launch_status();
    }
    if (!rc) rc = skk_timeline_mark(1);
    if (!rc) rc = skk_sync();
    if (!rc) {
        const double ms = skk_timeline_ms(0, 1) / reps;
        *rate = ms > 0.0 ? (double)count / (ms * 1e-3) : 0.0;
    }
    skk_free(idx); skk_free(val); skk_free(tab); skk_free(out);
    return rc;
}

}  // extern "C"
