// Plasma charge/current deposit (plasma3d/weights.py:35-67, 249-285), gfx950.
//
// The reference scatters 25 taps x 4 grids per particle (serially on the
// CPU, with float64 global atomics on the GPU).  Here the deposit is a
// gather over cell-sorted particles:
//
//   1. bin      cell of every particle (fused into the push when possible),
//               integer atomic count per cell;
//   2. scan     exclusive prefix of the counts -> start of each cell's list;
//   3. scatter  particle ids into their cell's list (the integer counters
//               are counted back down to zero, ready for the next deposit);
//   4. order    each cell's list sorted by particle id, so the summation
//               order below does not depend on atomic arrival order;
//   5. gather   one workgroup per 16x16 tile of output cells stages the
//               particles of the 20x20 source cells around it in LDS
//               (quartic weights / h and the four charges per particle),
//               then each thread sums, for its own output cell, the
//               contributions of the 5 source rows in ascending order.
//
// Every output cell is written by exactly one thread: no floating-point
// atomics, and the result is bit-reproducible run to run.  In the FP32-bulk
// form the tile accumulators are Kahan-compensated.
#include "common.h"
#include "kernels.h"

namespace lc3d {

constexpr int DEP_TILE = 16;
constexpr int DEP_HALO = DEP_TILE + 4;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER;
constexpr int MAX_SORTED_CELL = 4096;

// ---- binning ------------------------------------------------------------------
template <typename TP>
__global__ __launch_bounds__(256) void k_bin(int64_t n, double h, int N,
                                             const double *__restrict__ x_init,
                                             const double *__restrict__ y_init,
                                             const double *__restrict__ xo,
                                             const double *__restrict__ yo,
                                             const TP *__restrict__ q, BinState bs)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    int key = -1;
    if (q[k] != TP(0)) {
        int ix, iy;
        double xl, yl;
        bool finite;
        if (stencil_cell(x_init[k] + xo[k], y_init[k] + yo[k], h, N, ix, iy, xl, yl, finite)) {
            key = ix * N + iy;
            atomicAdd(&bs.count[key], 1);
        } else {
            atomicOr(bs.err, finite ? ERR_PARTICLE_OFF_GRID : ERR_NONFINITE);
        }
    }
    bs.key[k] = key;
}

// ---- exclusive scan of the N*N cell counts into start[0..N*N] --------------------
__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total)
{
    // Inclusive scan over the 256 threads (Hillis-Steele in LDS).
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const int add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[SCAN_THREADS - 1];
    const int incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(int ncell, const int *__restrict__ count,
                                                              int *__restrict__ blocks)
{
    __shared__ int sh[SCAN_THREADS];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i)
        if (base + i < ncell) s += count[base + i];
    int total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) blocks[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_final(int ncell, int nblocks,
                                                             const int *__restrict__ count,
                                                             const int *__restrict__ blocks,
                                                             int *__restrict__ start)
{
    __shared__ int sh[SCAN_THREADS];
    // Offset of this tile: sum of the block totals before it.
    int pre = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += SCAN_THREADS) pre += blocks[i];
    int dummy;
    const int my = block_exclusive_scan(pre, sh, dummy);
    (void)my;
    const int offset = dummy;   // total over all threads = sum of earlier blocks
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int c[SCAN_PER];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        c[i] = base + i < ncell ? count[base + i] : 0;
        s += c[i];
    }
    int total;
    int run = offset + block_exclusive_scan(s, sh, total);
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        if (base + i <= ncell) start[base + i] = run;
        run += c[i];
    }
    (void)nblocks;
}

// ---- scatter ids into the cell lists, counting the counters back to zero --------
__global__ __launch_bounds__(256) void k_scatter(int64_t n, BinState bs)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int key = bs.key[k];
    if (key < 0) return;
    const int slot = bs.start[key] + atomicSub(&bs.count[key], 1) - 1;
    bs.idx[slot] = (int)k;
}

// ---- sort each cell's list by particle id ---------------------------------------
__global__ __launch_bounds__(256) void k_order_cells(int ncell, BinState bs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int s = bs.start[c], e = bs.start[c + 1];
    if (e - s < 2) return;
    if (e - s > MAX_SORTED_CELL) {   // degenerate pile-up: sum in arrival order
        atomicOr(bs.err, ERR_CELL_OVERFLOW);
        return;
    }
    int *a = bs.idx;
    for (int i = s + 1; i < e; ++i) {
        const int v = a[i];
        int j = i - 1;
        while (j >= s && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
}

// ---- gather deposit -------------------------------------------------------------------
template <typename T> struct GatherCap;
template <> struct GatherCap<double> { static constexpr int value = 512; };
template <> struct GatherCap<float> { static constexpr int value = 1024; };

template <typename T, bool KAHAN> struct Acc;
template <typename T> struct Acc<T, false> {
    T s = 0;
    __device__ __forceinline__ void add(T v) { s += v; }
    __device__ __forceinline__ T get() const { return s; }
};
template <typename T> struct Acc<T, true> {
    T s = 0, c = 0;
    __device__ __forceinline__ void add(T v) {
        const T y = v - c;
        const T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ T get() const { return s; }
};

template <typename TP, typename TG, bool KAHAN>
__global__ __launch_bounds__(256) void k_deposit_gather(
    int N, double h, int tiles, const int *__restrict__ start, const int *__restrict__ idx,
    const double *__restrict__ x_init, const double *__restrict__ y_init,
    const double *__restrict__ xo, const double *__restrict__ yo, const TP *__restrict__ px,
    const TP *__restrict__ py, const TP *__restrict__ pz, const TP *__restrict__ q,
    const TP *__restrict__ m, TG *__restrict__ ro, TG *__restrict__ jx, TG *__restrict__ jy,
    TG *__restrict__ jz, const TG *__restrict__ bg_ro, TG *__restrict__ ro1,
    TG *__restrict__ jx1, TG *__restrict__ jy1, TG *__restrict__ jz1)
{
    using T = TP;   // arithmetic type of the tile
    constexpr int CAP = GatherCap<T>::value;
    __shared__ int hs[DEP_HALO][DEP_HALO + 1];   // global list position of each halo cell
    __shared__ int rowoff[DEP_HALO + 1];         // halo-list offset of each halo row
    __shared__ T s_wx[5][CAP];
    __shared__ T s_wy[5][CAP];
    __shared__ T s_c[4][CAP];
    __shared__ int s_cy[CAP];

    const int tile = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int X0 = (tile / tpr) * DEP_TILE, Y0 = (tile % tpr) * DEP_TILE;
    (void)tiles;
    const int hx0 = max(0, X0 - 2), hx1 = min(N, X0 + DEP_TILE + 2);
    const int hy0 = max(0, Y0 - 2), hy1 = min(N, Y0 + DEP_TILE + 2);
    const int nrow = hx1 - hx0, ncol = hy1 - hy0;
    const int t = threadIdx.x;

    for (int i = t; i < nrow * (ncol + 1); i += 256) {
        const int r = i / (ncol + 1), c = i % (ncol + 1);
        hs[r][c] = start[(hx0 + r) * N + hy0 + c];
    }
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int r = 0; r < nrow; ++r) {
            rowoff[r] = acc;
            acc += hs[r][ncol] - hs[r][0];
        }
        rowoff[nrow] = acc;
    }
    __syncthreads();
    const int total = rowoff[nrow];

    // This thread's output cell and, per source row, its particle range.
    const int tx = t / DEP_TILE, ty = t % DEP_TILE;
    const int ox = X0 + tx, oy = Y0 + ty;
    const bool valid = ox < N && oy < N;
    int lo[5], hi[5];
#pragma unroll
    for (int kr = 0; kr < 5; ++kr) {
        const int r = ox - 2 + kr - hx0;
        lo[kr] = hi[kr] = 0;
        if (valid && r >= 0 && r < nrow) {
            const int c0 = max(oy - 2, hy0) - hy0, c1 = min(oy + 3, hy1) - hy0;
            lo[kr] = rowoff[r] + hs[r][c0] - hs[r][0];
            hi[kr] = rowoff[r] + hs[r][c1] - hs[r][0];
        }
    }

    Acc<T, KAHAN> a_ro, a_jx, a_jy, a_jz;
    const T inv_h = T(1.0 / h);
    for (int k0 = 0; k0 < total; k0 += CAP) {
        const int k1 = min(total, k0 + CAP);
        // Stage: quartic weights (pre-divided by h) and charges per particle.
        for (int j = k0 + t; j < k1; j += 256) {
            int r = 0;
            while (rowoff[r + 1] <= j) ++r;
            const int p = idx[hs[r][0] + (j - rowoff[r])];
            int ix, iy;
            double xl, yl;
            bool finite;
            stencil_cell(x_init[p] + xo[p], y_init[p] + yo[p], h, N, ix, iy, xl, yl, finite);
            T w[5];
            weights4<T>((T)xl, w);
#pragma unroll
            for (int i = 0; i < 5; ++i) s_wx[i][j - k0] = w[i] * inv_h;
            weights4<T>((T)yl, w);
#pragma unroll
            for (int i = 0; i < 5; ++i) s_wy[i][j - k0] = w[i] * inv_h;
            const double P = px[p], Q = py[p], R = pz[p], M = m[p];
            const double gm = sqrt(M * M + P * P + Q * Q + R * R);
            const double dro = (double)q[p] / (1 - R / gm);
            const double r_g = dro / gm;
            s_c[0][j - k0] = (T)dro;
            s_c[1][j - k0] = (T)(P * r_g);
            s_c[2][j - k0] = (T)(Q * r_g);
            s_c[3][j - k0] = (T)(R * r_g);
            s_cy[j - k0] = iy;
        }
        __syncthreads();
        // Accumulate source rows in ascending order: kx = ox - sx = 2 - kr.
#pragma unroll
        for (int kr = 0; kr < 5; ++kr) {
            const int b = max(lo[kr], k0), e = min(hi[kr], k1);
            for (int j = b; j < e; ++j) {
                const int jj = j - k0;
                const T w = s_wx[4 - kr][jj] * s_wy[oy - s_cy[jj] + 2][jj];
                a_ro.add(s_c[0][jj] * w);
                a_jx.add(s_c[1][jj] * w);
                a_jy.add(s_c[2][jj] * w);
                a_jz.add(s_c[3][jj] * w);
            }
        }
        __syncthreads();
    }

    if (valid) {
        const int o = ox * N + oy;
        ro[o] = (TG)a_ro.get();
        jx[o] = (TG)a_jx.get();
        jy[o] = (TG)a_jy.get();
        jz[o] = (TG)a_jz.get();
        if (bg_ro != nullptr) {
            ro1[o] = bg_ro[o];
            jx1[o] = 0;
            jy1[o] = 0;
            jz1[o] = 0;
        }
    }
}

template <typename TP, typename TG>
hipError_t launch_deposit(const PushGeom &g, const SortArrays<TP> &a, bool use_full, bool binned,
                          const BinState &bs, TG *ro, TG *jx, TG *jy, TG *jz, const TG *bg_ro,
                          TG *ro1, TG *jx1, TG *jy1, TG *jz1, hipStream_t s)
{
    const int N = g.N;
    const int ncell = N * N;
    const double *xo = use_full ? a.xo : a.xo_prev;
    const double *yo = use_full ? a.yo : a.yo_prev;
    const TP *px = use_full ? a.px : a.px_prev;
    const TP *py = use_full ? a.py : a.py_prev;
    const TP *pz = use_full ? a.pz : a.pz_prev;
    const TP *q = use_full ? a.q : a.q_prev;
    const int pblocks = (int)((g.n + 255) / 256);
    if (g.n > 0) {
        if (!binned)
            k_bin<TP><<<pblocks, 256, 0, s>>>(g.n, g.h, N, a.x_init, a.y_init, xo, yo, q, bs);
        const int nb = (ncell + 1 + SCAN_TILE - 1) / SCAN_TILE;
        k_scan_blocks<<<nb, SCAN_THREADS, 0, s>>>(ncell, bs.count, bs.blocks);
        k_scan_final<<<nb, SCAN_THREADS, 0, s>>>(ncell, nb, bs.count, bs.blocks, bs.start);
        k_scatter<<<pblocks, 256, 0, s>>>(g.n, bs);
        k_order_cells<<<(ncell + 255) / 256, 256, 0, s>>>(ncell, bs);
    } else {
        // Empty sort: every cell list is empty.
        const int nb = (ncell + 1 + SCAN_TILE - 1) / SCAN_TILE;
        k_scan_blocks<<<nb, SCAN_THREADS, 0, s>>>(ncell, bs.count, bs.blocks);
        k_scan_final<<<nb, SCAN_THREADS, 0, s>>>(ncell, nb, bs.count, bs.blocks, bs.start);
    }
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int tiles = tpr * tpr;
    constexpr bool kahan = sizeof(TP) == 4;
    k_deposit_gather<TP, TG, kahan><<<tiles, 256, 0, s>>>(
        N, g.h, tiles, bs.start, bs.idx, a.x_init, a.y_init, xo, yo, px, py, pz, q, a.m, ro, jx,
        jy, jz, bg_ro, ro1, jx1, jy1, jz1);
    return hipGetLastError();
}

template hipError_t launch_deposit<double, double>(const PushGeom &, const SortArrays<double> &,
                                                   bool, bool, const BinState &, double *,
                                                   double *, double *, double *, const double *,
                                                   double *, double *, double *, double *,
                                                   hipStream_t);
template hipError_t launch_deposit<float, float>(const PushGeom &, const SortArrays<float> &,
                                                 bool, bool, const BinState &, float *, float *,
                                                 float *, float *, const float *, float *,
                                                 float *, float *, float *, hipStream_t);

}  // namespace lc3d
