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

#include <type_traits>

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
                                             const TP *__restrict__ px, const TP *__restrict__ py,
                                             const TP *__restrict__ pz, const TP *__restrict__ q,
                                             const TP *__restrict__ m, BinState bs)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    int key = -1;
    if (q[k] != TP(0)) {
        key = bin_particle<TP>(x_init[k] + xo[k], y_init[k] + yo[k], h, N, (double)px[k],
                               (double)py[k], (double)pz[k], (double)m[k], (double)q[k],
                               static_cast<DepRec<TP> *>(bs.rec), k, bs.err);
        if (key >= 0) atomicAdd(&bs.count[key], 1);
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
// Particle slots per halo cell staged at once.
template <typename T> struct SlotCap;
template <> struct SlotCap<double> { static constexpr int value = 2; };
template <> struct SlotCap<float> { static constexpr int value = 4; };

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

// One block per 16 x 16 tile of output cells, one thread per output cell.
// The records (written by the pass that binned each particle) of the 20 x 20
// halo cells' particles are staged into LDS slot by slot
// ([slot][halo row][halo column]), so the 64 lanes of a wave read 16
// consecutive cells of 4 rows and all run the same 25 x slots loop.  Each
// output cell sums its 5 x 5 source cells row by row (x ascending), then by
// column, then in particle-id order within a cell; passes over slots K at a
// time interleave this order when a halo cell holds more than K particles.
// No floating-point atomics: the result is reproducible bit for bit.
template <typename TP, typename TG, bool KAHAN>
__global__ __launch_bounds__(256) void k_deposit_gather(
    int N, double h, const int *__restrict__ start, BinState bs, TG *__restrict__ ro,
    TG *__restrict__ jx, TG *__restrict__ jy, TG *__restrict__ jz, const TG *__restrict__ bg_ro,
    TG *__restrict__ ro1, TG *__restrict__ jx1, TG *__restrict__ jy1, TG *__restrict__ jz1)
{
    using T = TP;   // arithmetic type of the tile
    constexpr int K = SlotCap<T>::value;
    constexpr int H = DEP_HALO;
    __shared__ int s_start[H][H];
    __shared__ int s_cnt[H][H];
    __shared__ T s_xl[K][H][H];
    __shared__ T s_yl[K][H][H];
    __shared__ T s_c[4][K][H][H];
    __shared__ int s_max;

    const int tile = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int X0 = (tile / tpr) * DEP_TILE, Y0 = (tile % tpr) * DEP_TILE;
    const DepRec<T> *__restrict__ rec = static_cast<const DepRec<T> *>(bs.rec);
    const int t = threadIdx.x;
    if (t == 0) s_max = 0;
    __syncthreads();
    int my_max = 0;
    for (int i = t; i < H * H; i += 256) {
        const int r = i / H, c = i % H;
        const int gx = X0 - 2 + r, gy = Y0 - 2 + c;
        int st = 0, cnt = 0;
        if (gx >= 0 && gx < N && gy >= 0 && gy < N) {
            const int cell = gx * N + gy;
            st = start[cell];
            cnt = start[cell + 1] - st;
        }
        s_start[r][c] = st;
        s_cnt[r][c] = cnt;
        my_max = max(my_max, cnt);
    }
    atomicMax(&s_max, my_max);
    __syncthreads();
    const int maxcnt = s_max;

    const int tx = t / DEP_TILE, ty = t % DEP_TILE;
    const int ox = X0 + tx, oy = Y0 + ty;
    Acc<T, KAHAN> a_ro, a_jx, a_jy, a_jz;
    const T inv_h = T(1.0 / h);
    for (int k0 = 0; k0 < maxcnt; k0 += K) {
        for (int e = t; e < K * H * H; e += 256) {
            const int k = e / (H * H), cell = e % (H * H);
            const int r = cell / H, c = cell % H;
            T xl = 0, yl = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
            if (k0 + k < s_cnt[r][c]) {
                const DepRec<T> d = rec[bs.idx[s_start[r][c] + k0 + k]];
                xl = d.xl;
                yl = d.yl;
                c0 = d.c0;
                c1 = d.c1;
                c2 = d.c2;
                c3 = d.c3;
            }
            s_xl[k][r][c] = xl;
            s_yl[k][r][c] = yl;
            s_c[0][k][r][c] = c0;
            s_c[1][k][r][c] = c1;
            s_c[2][k][r][c] = c2;
            s_c[3][k][r][c] = c3;
        }
        __syncthreads();
        // Source cell (tx + sx, ty + sy) in halo coordinates lies at offset
        // (2 - sx, 2 - sy) from this output cell: taps 4 - sx and 4 - sy.
        const int kn = min(K, maxcnt - k0);
        auto row = [&](auto sxc) {
            constexpr int sx = decltype(sxc)::value;
            auto col = [&](auto syc) {
                constexpr int sy = decltype(syc)::value;
                for (int k = 0; k < kn; ++k) {
                    const T wx = w4_tap<4 - sx>(s_xl[k][tx + sx][ty + sy]) * inv_h;
                    const T wy = w4_tap<4 - sy>(s_yl[k][tx + sx][ty + sy]) * inv_h;
                    const T w = wx * wy;
                    a_ro.add(s_c[0][k][tx + sx][ty + sy] * w);
                    a_jx.add(s_c[1][k][tx + sx][ty + sy] * w);
                    a_jy.add(s_c[2][k][tx + sx][ty + sy] * w);
                    a_jz.add(s_c[3][k][tx + sx][ty + sy] * w);
                }
            };
            col(std::integral_constant<int, 0>{});
            col(std::integral_constant<int, 1>{});
            col(std::integral_constant<int, 2>{});
            col(std::integral_constant<int, 3>{});
            col(std::integral_constant<int, 4>{});
        };
        row(std::integral_constant<int, 0>{});
        row(std::integral_constant<int, 1>{});
        row(std::integral_constant<int, 2>{});
        row(std::integral_constant<int, 3>{});
        row(std::integral_constant<int, 4>{});
        __syncthreads();
    }

    if (ox < N && oy < N) {
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
            k_bin<TP><<<pblocks, 256, 0, s>>>(g.n, g.h, N, a.x_init, a.y_init, xo, yo, px, py, pz,
                                              q, a.m, bs);
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
    k_deposit_gather<TP, TG, kahan><<<tiles, 256, 0, s>>>(N, g.h, bs.start, bs, ro, jx, jy, jz,
                                                          bg_ro, ro1, jx1, jy1, jz1);
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
