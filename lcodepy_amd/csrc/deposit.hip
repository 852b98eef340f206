// Plasma charge/current deposit (plasma3d/weights.py:35-67, 249-285), gfx950.
//
// The reference scatters 25 taps x 4 grids per particle (serially on the
// CPU, with float64 global atomics on the GPU).  Here the deposit is a
// gather over cell-sorted particles:
//
//   1. bin      cell of every particle (fused into the push when possible),
//               integer atomic count per cell (one atomic per run of equal
//               cells in a wave), which also gives each particle its rank in
//               its cell;
//   2. scan     exclusive prefix of the counts -> start of each cell's list
//               (the counts are cleared for the next deposit);
//   3. scatter  particle ids into their cell's list at their ranks;
//   4. order    each cell's list sorted by particle id (by the cell's thread,
//               or for a dense cell by a workgroup in step 5), so the
//               summation order below does not depend on atomic arrival
//               order;
//   5. taps     per dense cell (more than 16 particles) its list sorted and
//               its 5 x 5 x 4 taps, by one workgroup with a fixed-order
//               reduction;
//   6. gather   one workgroup per 16x16 tile of output cells forms, in five
//               passes over the x offset, the taps of the 20x20 source
//               cells around it in LDS; each thread adds, for its own output
//               cell, the 25 taps that reach it in a fixed order.
//
// Every output cell is written by exactly one thread: no floating-point
// atomics, and the result is bit-reproducible run to run.  In the FP32-bulk
// form the tile accumulators are Kahan-compensated.
#include "common.h"
#include "kernels.h"

#include <climits>
#include <type_traits>

namespace lc3d {

constexpr int DEP_TILE = 16;
constexpr int DEP_HALO = DEP_TILE + 4;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER;
constexpr int SMALL_CELL = 16;          // lists sorted by their cell's thread
constexpr int MAX_SORTED_CELL = 16384;  // lists sorted by a workgroup in LDS

// ---- binning ------------------------------------------------------------------
template <typename TP>
__global__ __launch_bounds__(256) void k_bin(int64_t n, double h, int N,
                                             const double *__restrict__ x_init,
                                             const double *__restrict__ y_init,
                                             const double *__restrict__ xo,
                                             const double *__restrict__ yo,
                                             const TP *__restrict__ px, const TP *__restrict__ py,
                                             const double *__restrict__ pz, const TP *__restrict__ q,
                                             const TP *__restrict__ m, BinState bs)
{
    // Threads past the end stay to the count so that whole waves reach it.
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int key = -1;
    if (k < n) {
        if (q[k] != TP(0))
            key = bin_particle<TP>(x_init[k] + xo[k], y_init[k] + yo[k], h, N, (double)px[k],
                                   (double)py[k], (double)pz[k], (double)m[k], (double)q[k],
                                   static_cast<DepRec<TP> *>(bs.rec), k, bs.err);
        bs.key[k] = key;
    }
    // Unbinned particles are counted (and ranked) in count[N * N].
    const int r = count_rank(bs.count, k < n ? (key >= 0 ? key : N * N) : -1);
    if (k < n) bs.rank[k] = r;
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

// Writes start[0..ncell] and clears the counts it read (ready for the next
// binning pass).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_final(int ncell, int nblocks,
                                                             int *__restrict__ count,
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
        if (base + i < ncell) count[base + i] = 0;
        run += c[i];
    }
    (void)nblocks;
}

// ---- scatter ids into the cell lists --------------------------------------------
// Places particle k at its rank in its cell's list; particles that are not
// binned (q = 0 or off the grid) go after all the lists, at their rank among
// the unbinned ones, so that idx is a permutation of the sort.  count[ncell]
// counts them and is cleared by k_order_cells.
__global__ __launch_bounds__(256) void k_scatter(int64_t n, int ncell, BinState bs)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (k == 0) *bs.nbig = 0;   // k_order_cells refills the big-cell list
    const int key = bs.key[k];
    bs.idx[bs.start[key < 0 ? ncell : key] + bs.rank[k]] = (int)k;
}

// ---- sort each cell's list by particle id ---------------------------------------
// Lists of up to SMALL_CELL particles are sorted by their cell's thread
// (insertion sort).  Longer ones make the cell dense: it is queued (its slot
// recorded in dslot) for k_dense_cells, which sorts it.
__global__ __launch_bounds__(256) void k_order_cells(int ncell, BinState bs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) bs.count[ncell] = 0;   // the scatter's count of unbinned particles
    if (c >= ncell) return;
    const int s = bs.start[c], e = bs.start[c + 1];
    int slot = -1;
    if (e - s > SMALL_CELL) {
        // At most n / (SMALL_CELL + 1) < bigcap cells can be dense.
        slot = atomicAdd(bs.nbig, 1);
        if (slot < bs.bigcap) bs.big[slot] = c;
        else slot = -1;
        if (e - s > MAX_SORTED_CELL) atomicOr(bs.err, ERR_CELL_OVERFLOW);
    }
    bs.dslot[c] = slot;
    if (slot >= 0) return;
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

// The deposit records in list order: rec_sorted[i] = rec[idx[i]] for every
// binned list position i < start[ncell], one thread per 16-byte word of a
// record (a record is rec_words words), so the writes are contiguous.  The
// gather reads the non-dense cells' records from here.
__global__ __launch_bounds__(256) void k_sort_records(int ncell, BinState bs, int rec_words)
{
    const int64_t nwords = (int64_t)bs.start[ncell] * rec_words;
    const int4 *__restrict__ src = static_cast<const int4 *>(bs.rec);
    int4 *__restrict__ dst = static_cast<int4 *>(bs.rec_sorted);
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < nwords;
         o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = o / rec_words;
        const int w = (int)(o - i * rec_words);
        dst[o] = src[(int64_t)bs.idx[i] * rec_words + w];
    }
}

// ---- dense cells: sort and taps ------------------------------------------------------
// One workgroup per dense cell, the grid striding over the queue.  The cell's
// list is sorted by particle id with a bitonic sort in LDS (lists of up to
// MAX_SORTED_CELL; longer ones keep arrival order) and written back, so the
// next push visits the cell's particles in id order.  Then thread t sums its
// contiguous segment of the sorted list into the 5 x 5 taps x 4 currents
// around the cell; the 256 partial sums are combined by a butterfly within
// each wave and then the four waves in a fixed order.  out[slot][tap][q],
// tap = 5 a + b for the target cell (ix + a - 2, iy + b - 2), q = ro jx jy jz.
constexpr int DENSE_VALUES = 25 * 4;
constexpr int DENSE_THREADS = 256;
template <typename T>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense_cells(double h, BinState bs)
{
    __shared__ int s[MAX_SORTED_CELL];
    __shared__ T red[4][DENSE_VALUES];
    const int nbig = min(*bs.nbig, bs.bigcap);
    const DepRec<T> *__restrict__ rec = static_cast<const DepRec<T> *>(bs.rec);
    T *__restrict__ out = static_cast<T *>(bs.dense);
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    const T inv_h = T(1.0 / h);
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int c = bs.big[b];
        const int st = bs.start[c], n = bs.start[c + 1] - st;   // n > SMALL_CELL
        const bool sorted = n <= MAX_SORTED_CELL;
        if (sorted) {
            int P = 1;
            while (P < n) P <<= 1;
            for (int i = t; i < P; i += DENSE_THREADS) s[i] = i < n ? bs.idx[st + i] : INT_MAX;
            __syncthreads();
            for (int k = 2; k <= P; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = t; i < P; i += DENSE_THREADS) {
                        const int l = i ^ j;
                        if (l > i) {
                            const int x = s[i], y = s[l];
                            if ((x > y) == ((i & k) == 0)) {
                                s[i] = y;
                                s[l] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = t; i < n; i += DENSE_THREADS) bs.idx[st + i] = s[i];
        }
        const int seg = (n + DENSE_THREADS - 1) / DENSE_THREADS;
        const int j0 = min(n, t * seg), j1 = min(n, (t + 1) * seg);
        T acc[DENSE_VALUES];
#pragma unroll
        for (int k = 0; k < DENSE_VALUES; ++k) acc[k] = T(0);
        for (int j = j0; j < j1; ++j) {
            const DepRec<T> d = rec[sorted ? s[j] : bs.idx[st + j]];
            T wx[5], wy[5];
            weights4<T>(d.xl, wx);
            weights4<T>(d.yl, wy);
#pragma unroll
            for (int a = 0; a < 5; ++a) {
#pragma unroll
                for (int bb = 0; bb < 5; ++bb) {
                    const T w = (wx[a] * inv_h) * (wy[bb] * inv_h);
                    acc[(5 * a + bb) * 4 + 0] += d.c0 * w;
                    acc[(5 * a + bb) * 4 + 1] += d.c1 * w;
                    acc[(5 * a + bb) * 4 + 2] += d.c2 * w;
                    acc[(5 * a + bb) * 4 + 3] += d.c3 * w;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < DENSE_VALUES; ++k) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < DENSE_VALUES; ++k) red[wave][k] = acc[k];
        }
        __syncthreads();
        if (t < DENSE_VALUES)
            out[(size_t)b * DENSE_VALUES + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        __syncthreads();
    }
}

// ---- gather deposit -------------------------------------------------------------------
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

template <typename T> struct alignas(4 * sizeof(T)) Tap4 {
    T v[4];
};

// The five taps toward x offset A - 2 (target cells (ix + A - 2, iy + b - 2),
// b = 0..4) of one source cell, for the four currents: the sum over its
// particles, in list order, of c_q (w_A(xl) / h) (w_b(yl) / h), from the
// records in list order (rec = rec_sorted, first = the cell's list start),
// or the precomputed block of a dense cell.
// Records of a halo cell's first particles kept in registers for all five
// passes of the gather.
template <typename T> struct TapPrefetch {
    static constexpr int PF = sizeof(T) == 4 ? 4 : 2;
};
// Gather blocks per CU asked of the compiler: five in float32, whose 31 KB of
// LDS taps allow it, so that the 33 x 33 tiles of a 513 x 513 grid run in one
// round; two in float64 (62 KB).
template <typename T> struct GatherOcc {
    static constexpr int value = sizeof(T) == 4 ? 5 : 2;
};

template <int A, typename T>
__device__ __forceinline__ void add_taps(const DepRec<T> &d, T inv_h, Tap4<T> tp[5])
{
    const T wx = w4_tap<A>(d.xl) * inv_h;
    T wy[5];
    weights4<T>(d.yl, wy);
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const T w = wx * (wy[b] * inv_h);
        tp[b].v[0] += d.c0 * w;
        tp[b].v[1] += d.c1 * w;
        tp[b].v[2] += d.c2 * w;
        tp[b].v[3] += d.c3 * w;
    }
}

// The five taps toward x offset A - 2 (target cells (ix + A - 2, iy + b - 2),
// b = 0..4) of one source cell, for the four currents: the sum over its
// particles, in list order, of c_q (w_A(xl) / h) (w_b(yl) / h) -- the first
// PF (0 or TapPrefetch) records from the registers `pre`, the others from rec
// (= rec_sorted, first = the cell's list start) -- or the precomputed block
// of a dense cell.
template <int A, typename T, int PF>
__device__ __forceinline__ void cell_taps(const DepRec<T> *__restrict__ rec,
                                          const T *__restrict__ dense, int first, int n, int dsl,
                                          const DepRec<T> *pre, T inv_h, Tap4<T> tp[5])
{
#pragma unroll
    for (int b = 0; b < 5; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) tp[b].v[q] = T(0);
    if (dsl >= 0) {
        const T *v = dense + (size_t)dsl * DENSE_VALUES + 4 * (5 * A);
#pragma unroll
        for (int b = 0; b < 5; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) tp[b].v[q] = v[4 * b + q];
        return;
    }
#pragma unroll
    for (int k = 0; k < PF; ++k)
        if (k < n) add_taps<A, T>(pre[k], inv_h, tp);
    // The rest, up to four records in flight at a time.
    for (int j0 = PF; j0 < n; j0 += 4) {
        DepRec<T> d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (j0 + k < n) d[k] = rec[first + j0 + k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (j0 + k >= n) break;
            add_taps<A, T>(d[k], inv_h, tp);
        }
    }
}

// One block per 16 x 16 tile of output cells, one thread per output cell,
// in five passes over the x offset A of the source cells (source row
// ox - (A - 2)).  In a pass every cell of the tile's 20 x 20 halo forms its
// five taps toward that offset (cell_taps: from the deposit records its
// particles' binning pass wrote, in list order, or a dense cell's
// precomputed block) into LDS; then each output cell adds the five taps that
// reach it (source column oy - (b - 2), b ascending).  Every output cell is
// thus the same fixed-order sum of 25 taps, written by exactly one thread:
// no floating-point atomics, reproducible bit for bit.
template <typename TP, typename TG, bool KAHAN>
__global__ __launch_bounds__(256, GatherOcc<TP>::value) void k_deposit_gather(
    int N, double h, const int *__restrict__ start, BinState bs, TG *__restrict__ ro,
    TG *__restrict__ jx, TG *__restrict__ jy, TG *__restrict__ jz, const TG *__restrict__ bg_ro,
    TG *__restrict__ ro1, TG *__restrict__ jx1, TG *__restrict__ jy1, TG *__restrict__ jz1)
{
    using T = TP;   // arithmetic type of the taps
    constexpr int H = DEP_HALO;
    __shared__ Tap4<T> s_tap[5][H * H];

    const int tile = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int X0 = (tile / tpr) * DEP_TILE, Y0 = (tile % tpr) * DEP_TILE;
    const DepRec<T> *__restrict__ rec = static_cast<const DepRec<T> *>(bs.rec_sorted);
    const T *__restrict__ dense = static_cast<const T *>(bs.dense);
    const int hx0 = max(0, X0 - 2), hx1 = min(N, X0 + DEP_TILE + 2);
    const int hy0 = max(0, Y0 - 2), hy1 = min(N, Y0 + DEP_TILE + 2);
    const int nrow = hx1 - hx0, ncol = hy1 - hy0, nc = nrow * ncol;
    const int t = threadIdx.x;

    // This thread's halo cells, t and t + 256 (nc <= 400): list start,
    // particle count and dense slot (or -1).
    int c_first[2] = {0, 0}, c_cnt[2] = {0, 0}, c_dsl[2] = {-1, -1};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int i = t + 256 * u;
        if (i < nc) {
            const int cell = (hx0 + i / ncol) * N + hy0 + i % ncol;
            c_first[u] = start[cell];
            c_cnt[u] = start[cell + 1] - c_first[u];
            c_dsl[u] = bs.dslot[cell];
        }
    }

    const int tx = t / DEP_TILE, ty = t % DEP_TILE;
    const int ox = X0 + tx, oy = Y0 + ty;
    const bool valid = ox < N && oy < N;
    const T inv_h = T(1.0 / h);
    Acc<T, KAHAN> acc[4];
    static_assert(DEP_HALO * DEP_HALO <= 512, "two halo cells per thread");

    // This thread's halo cells (t and t + 256; nc <= 400) and their first
    // records, loaded once for the five passes.
    // (Only the first cell's: a thread's second cell, t + 256 < nc, reads its
    // records in each pass.)
    constexpr int PF = TapPrefetch<T>::PF;
    DepRec<T> pre[PF];
    if (t < nc && c_dsl[0] < 0) {
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (k < c_cnt[0]) pre[k] = rec[c_first[0] + k];
    }

    auto pass = [&](auto a_const) {
        constexpr int A = decltype(a_const)::value;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = t + 256 * u;
            if (i < nc) {
                Tap4<T> tp[5];
                if (u == 0)
                    cell_taps<A, T, TapPrefetch<T>::PF>(rec, dense, c_first[0], c_cnt[0], c_dsl[0],
                                                        pre, inv_h, tp);
                else
                    cell_taps<A, T, 0>(rec, dense, c_first[1], c_cnt[1], c_dsl[1], pre, inv_h, tp);
#pragma unroll
                for (int b = 0; b < 5; ++b) s_tap[b][i] = tp[b];
            }
        }
        __syncthreads();
        const int r = ox - (A - 2) - hx0;
        if (valid && r >= 0 && r < nrow) {
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                const int c = oy - (b - 2) - hy0;
                if (c >= 0 && c < ncol) {
                    const Tap4<T> v = s_tap[b][r * ncol + c];
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q].add(v.v[q]);
                }
            }
        }
        __syncthreads();
    };
    pass(std::integral_constant<int, 0>{});
    pass(std::integral_constant<int, 1>{});
    pass(std::integral_constant<int, 2>{});
    pass(std::integral_constant<int, 3>{});
    pass(std::integral_constant<int, 4>{});

    if (valid) {
        const int o = ox * N + oy;
        ro[o] = (TG)acc[0].get();
        jx[o] = (TG)acc[1].get();
        jy[o] = (TG)acc[2].get();
        jz[o] = (TG)acc[3].get();
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
                          TG *ro1, TG *jx1, TG *jy1, TG *jz1, hipStream_t s, KernelTimer *timer)
{
    const int N = g.N;
    const int ncell = N * N;
    const double *xo = use_full ? a.xo : a.xo_prev;
    const double *yo = use_full ? a.yo : a.yo_prev;
    const TP *px = use_full ? a.px : a.px_prev;
    const TP *py = use_full ? a.py : a.py_prev;
    const double *pz = use_full ? a.pz : a.pz_prev;
    const TP *q = use_full ? a.q : a.q_prev;
    const int pblocks = (int)((g.n + 255) / 256);
    const int rec_words = (int)(sizeof(DepRec<TP>) / 16);
    hipError_t e;
    if (timer && (e = timer->begin(KC_DEPOSIT, 1, s)) != hipSuccess) return e;
    if (g.n > 0) {
        if (!binned)
            k_bin<TP><<<pblocks, 256, 0, s>>>(g.n, g.h, N, a.x_init, a.y_init, xo, yo, px, py, pz,
                                              q, a.m, bs);
        const int nb = (ncell + 1 + SCAN_TILE - 1) / SCAN_TILE;
        k_scan_blocks<<<nb, SCAN_THREADS, 0, s>>>(ncell, bs.count, bs.blocks);
        k_scan_final<<<nb, SCAN_THREADS, 0, s>>>(ncell, nb, bs.count, bs.blocks, bs.start);
        k_scatter<<<pblocks, 256, 0, s>>>(g.n, ncell, bs);
        k_order_cells<<<(ncell + 255) / 256, 256, 0, s>>>(ncell, bs);
        k_sort_records<<<(int)std::min<int64_t>(8192, (g.n * rec_words + 255) / 256), 256, 0, s>>>(
            ncell, bs, rec_words);
        k_dense_cells<TP><<<512, DENSE_THREADS, 0, s>>>(g.h, bs);
    } else {
        // Empty sort: every cell list is empty (and no cell is dense).
        const int nb = (ncell + 1 + SCAN_TILE - 1) / SCAN_TILE;
        k_scan_blocks<<<nb, SCAN_THREADS, 0, s>>>(ncell, bs.count, bs.blocks);
        k_scan_final<<<nb, SCAN_THREADS, 0, s>>>(ncell, nb, bs.count, bs.blocks, bs.start);
        k_order_cells<<<(ncell + 255) / 256, 256, 0, s>>>(ncell, bs);
    }
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int tiles = tpr * tpr;
    constexpr bool kahan = sizeof(TP) == 4;
    k_deposit_gather<TP, TG, kahan><<<tiles, 256, 0, s>>>(N, g.h, bs.start, bs, ro, jx, jy, jz,
                                                          bg_ro, ro1, jx1, jy1, jz1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timer) return timer->end(s);
    return hipSuccess;
}

template hipError_t launch_deposit<double, double>(const PushGeom &, const SortArrays<double> &,
                                                   bool, bool, const BinState &, double *,
                                                   double *, double *, double *, const double *,
                                                   double *, double *, double *, double *,
                                                   hipStream_t, KernelTimer *);
template hipError_t launch_deposit<float, float>(const PushGeom &, const SortArrays<float> &,
                                                 bool, bool, const BinState &, float *, float *,
                                                 float *, float *, const float *, float *,
                                                 float *, float *, float *, hipStream_t, KernelTimer *);

}  // namespace lc3d
