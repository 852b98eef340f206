// Plasma charge/current deposit (plasma3d/weights.py:35-67, 249-285), gfx950.
//
// The reference scatters 25 taps x 4 grids per particle (serially on the
// CPU, with float64 global atomics on the GPU).  Here the deposit is a
// gather over cell-sorted particles:
//
//   1. bin      cell of every particle (fused into the push when possible),
//               written with its deposit record at its binning position
//               (the binning thread's index, so the writes are contiguous),
//               integer atomic count per cell (one atomic per run of equal
//               cells in a wave), which also gives each particle its rank in
//               its cell;
//   2. scan     exclusive prefix of the counts -> start of each cell's list
//               (the counts are cleared for the next deposit);
//   3. scatter  binning positions into their cell's list at their ranks;
//   4. order    each cell's list sorted by particle id (by the cell's
//               thread, or for a dense cell by a workgroup in step 5), so
//               the summation order below depends neither on atomic arrival
//               order nor on the order the particles were visited in; the
//               particle ids of the lists are the cell order the next push
//               visits (so binning positions follow the cell order, and a
//               cell's records lie close together);
//   5. taps     per dense cell (more than 16 particles) its list sorted and
//               its 5 x 5 x 4 taps, by one workgroup with a fixed-order
//               reduction;
//   6. gather   one workgroup per 16x16 tile of output cells forms the taps
//               of the 20x20 source cells around it (records read through
//               the lists), one x offset at a time in LDS; each output cell
//               adds the 25 taps that reach it in a fixed order.
//
// Every output value is written by exactly one thread: no floating-point
// atomics, and the result is bit-reproducible run to run.  The tile sums are
// float64 in both forms.
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
static_assert(SCAN_TILE == SCAN_TILE_CELLS, "scan tiles as allocated");
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
    // Particle k is binned at position k.
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int key = -1;
    if (k < n) {
        if (q[k] != TP(0))
            key = bin_particle<TP>(x_init[k] + xo[k], y_init[k] + yo[k], h, N, (double)px[k],
                                   (double)py[k], (double)pz[k], (double)m[k], (double)q[k],
                                   static_cast<DepRec<TP> *>(bs.rec), k, bs.err);
        bs.key[k] = key;
        bs.pid[k] = (int)k;
    }
    // Unbinned particles are counted (and ranked) in count[N * N].
    const int r = count_rank(bs.count, k < n ? (key >= 0 ? key : N * N) : -1);
    if (k < n) bs.rank[k] = r;
}

// ---- exclusive scan of the N*N cell counts -----------------------------------------
// Inclusive scan of v over the wave's lanes in order (__shfl_up steps).
__device__ __forceinline__ int wave_inclusive_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d);
        if (lane >= d) v += u;
    }
    return v;
}

// Exclusive scan of v over the block's SCAN_THREADS threads (every thread
// calls it); total = the block's sum.  The waves scan themselves, then add
// the totals of the waves before them (sh: SCAN_THREADS / 64 ints).
__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total)
{
    constexpr int W = SCAN_THREADS / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int incl = wave_inclusive_scan(v);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    int off = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int t = sh[w];
        off += w < wave ? t : 0;
        sum += t;
    }
    total = sum;
    __syncthreads();   // sh is reused by the next call
    return off + incl - v;
}

// One pass, no block waiting for another.  Each block scans its tile of
// SCAN_TILE cells into lstart (starts within the tile, lstart[ncell] included)
// and clears the counts it read (ready for the next binning pass); then adds
// the tile's total to base[t] of every later tile t by integer atomics, so
// that when the pass has ended base[t] is the sum of the tiles before t,
// whatever order the blocks ran in.  A cell's list starts at
// lstart[c] + base[c / SCAN_TILE]: k_scatter reads it so, and k_order_cells
// writes it to start[].  base is zero when the pass begins (the gather clears
// it once the list sort has read it).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_cells(int ncell, int ntiles,
                                                             int *__restrict__ count,
                                                             int *__restrict__ lstart,
                                                             int *__restrict__ base)
{
    __shared__ int sh[SCAN_THREADS / 64];
    const int first = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int c[SCAN_PER];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        c[i] = first + i < ncell ? count[first + i] : 0;
        s += c[i];
    }
    int total;
    int run = block_exclusive_scan(s, sh, total);
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        if (first + i <= ncell) lstart[first + i] = run;
        if (first + i < ncell) count[first + i] = 0;
        run += c[i];
    }
    if (total != 0)
        for (int t = blockIdx.x + 1 + threadIdx.x; t < ntiles; t += SCAN_THREADS)
            atomicAdd(&base[t], total);
}

// Start of cell c's list (c <= ncell) after k_scan_cells.
__device__ __forceinline__ int list_start(const BinState &bs, int c)
{
    return bs.lstart[c] + bs.base[c / SCAN_TILE];
}

// ---- scatter positions into the cell lists ---------------------------------------
// Places binning position j, and the particle binned there, at its rank in
// its cell's list; particles that are not binned (q = 0 or off the grid) go
// after all the lists, at their rank among the unbinned ones, so that pos and
// idx are permutations.  count[ncell] counts them and is cleared by
// k_order_cells.
__global__ __launch_bounds__(256) void k_scatter(int64_t n, int ncell, BinState bs)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (j == 0) *bs.nbig = 0;   // k_order_cells refills the big-cell list
    const int key = bs.key[j];
    const int slot = list_start(bs, key < 0 ? ncell : key) + bs.rank[j];
    bs.pos[slot] = (int)j;
    bs.idx[slot] = bs.pid[j];
}

// ---- sort each cell's list by particle id ------------------------------------------
// Insertion sort of list [s, e) by particle id (idx), pos alongside, in memory.
__device__ void sort_list_in_memory(const BinState &bs, int s, int e)
{
    int *a = bs.pos, *d = bs.idx;
    for (int i = s + 1; i < e; ++i) {
        const int v = d[i], p = a[i];
        int j = i - 1;
        while (j >= s && d[j] > v) {
            d[j + 1] = d[j];
            a[j + 1] = a[j];
            --j;
        }
        d[j + 1] = v;
        a[j + 1] = p;
    }
}

// One thread per cell.  A list of up to SMALL_CELL particles has its
// particle ids read into registers and, if they are out of order (the usual
// list, whose particles were visited in the last deposit's order, is in order
// unless a particle arrived from another wave), is sorted in this thread's
// row of LDS and written back, pos alongside.  A longer list makes the cell
// dense: it is queued (its slot recorded in dslot) for k_dense_cells, which
// sorts it, or, when the queue is full, sorted here in memory.
__global__ __launch_bounds__(256) void k_order_cells(int ncell, BinState bs)
{
    __shared__ int s_d[256][SMALL_CELL + 1], s_a[256][SMALL_CELL + 1];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) bs.count[ncell] = 0;   // the scatter's count of unbinned particles
    if (c >= ncell) return;
    const int s = list_start(bs, c), e = list_start(bs, c + 1), n = e - s;
    bs.start[c] = s;
    if (c == ncell - 1) bs.start[ncell] = e;
    int slot = -1;
    if (n > SMALL_CELL) {
        // At most n / (SMALL_CELL + 1) < bigcap cells can be dense.
        slot = atomicAdd(bs.nbig, 1);
        if (slot < bs.bigcap) bs.big[slot] = c;
        else slot = -1;
        if (n > MAX_SORTED_CELL) atomicOr(bs.err, ERR_CELL_OVERFLOW);
        if (slot < 0) sort_list_in_memory(bs, s, e);
    }
    bs.dslot[c] = slot;
    if (n > SMALL_CELL) return;
    int d[SMALL_CELL];
#pragma unroll
    for (int k = 0; k < SMALL_CELL; ++k) d[k] = k < n ? bs.idx[s + k] : INT_MAX;
    bool sorted = true;
#pragma unroll
    for (int k = 1; k < SMALL_CELL; ++k)
        if (d[k - 1] > d[k]) sorted = false;   // the INT_MAX padding stays in order
    if (sorted) return;
    int *dd = s_d[threadIdx.x], *aa = s_a[threadIdx.x];
#pragma unroll
    for (int k = 0; k < SMALL_CELL; ++k) {
        if (k < n) {
            dd[k] = d[k];
            aa[k] = bs.pos[s + k];
        }
    }
    for (int i = 1; i < n; ++i) {
        const int v = dd[i], p = aa[i];
        int j = i - 1;
        while (j >= 0 && dd[j] > v) {
            dd[j + 1] = dd[j];
            aa[j + 1] = aa[j];
            --j;
        }
        dd[j + 1] = v;
        aa[j + 1] = p;
    }
    for (int k = 0; k < n; ++k) {
        bs.idx[s + k] = dd[k];
        bs.pos[s + k] = aa[k];
    }
}

// ---- dense cells: sort and taps ------------------------------------------------------
// One workgroup per dense cell, the grid striding over the queue.  The cell's
// list is sorted by particle id with a bitonic sort of (particle id,
// position) keys in LDS (lists of up to MAX_SORTED_CELL; longer ones keep
// arrival order) and written back, positions to pos and particle ids to idx
// (the next push's order).  Then thread t sums its
// contiguous segment of the sorted list into the 5 x 5 taps x 4 currents
// around the cell; the 256 partial sums are combined by a butterfly within
// each wave and then the four waves in a fixed order.  out[slot][tap][q],
// tap = 5 a + b for the target cell (ix + a - 2, iy + b - 2), q = ro jx jy jz.
constexpr int DENSE_VALUES = 25 * 4;
constexpr int DENSE_THREADS = 256;
template <typename T>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense_cells(double h, BinState bs)
{
    __shared__ unsigned long long s[MAX_SORTED_CELL];   // particle id << 32 | position
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
            for (int i = t; i < P; i += DENSE_THREADS) {
                unsigned long long v = ~0ull;
                if (i < n)
                    v = (unsigned long long)(unsigned)bs.idx[st + i] << 32 | (unsigned)bs.pos[st + i];
                s[i] = v;
            }
            __syncthreads();
            for (int k = 2; k <= P; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = t; i < P; i += DENSE_THREADS) {
                        const int l = i ^ j;
                        if (l > i) {
                            const unsigned long long x = s[i], y = s[l];
                            if ((x > y) == ((i & k) == 0)) {
                                s[i] = y;
                                s[l] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = t; i < n; i += DENSE_THREADS) {
                bs.pos[st + i] = (int)(unsigned)s[i];
                bs.idx[st + i] = (int)(s[i] >> 32);
            }
        }
        const int seg = (n + DENSE_THREADS - 1) / DENSE_THREADS;
        const int j0 = min(n, t * seg), j1 = min(n, (t + 1) * seg);
        T acc[DENSE_VALUES];
#pragma unroll
        for (int k = 0; k < DENSE_VALUES; ++k) acc[k] = T(0);
        for (int j = j0; j < j1; ++j) {
            const DepRec<T> d = rec[sorted ? (int)(unsigned)s[j] : bs.pos[st + j]];
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
template <typename T> struct alignas(4 * sizeof(T)) Tap4 {
    T v[4];
};

// One block per 16 x 16 tile of output cells.  Its 20 x 20 halo of source
// cells is dealt to the block's threads one cell each, ordered by particle
// count (so that the cells of a wave have about as many particles and the
// wave's loop over them wastes few lanes); a thread forms its cell's taps,
// group by group of x offsets, holding a group's taps in registers while it
// reads the cell's records once for the whole group.  Tap groups (number of
// x offsets A per group) and the waves per SIMD asked of the compiler
// (__launch_bounds__'s second argument, which bounds the registers):
// float32 taps in groups of 3 and 2 (60 values in registers), float64 one x
// offset at a time in a runtime loop (fewer registers, two blocks per CU);
// rf records of a cell in flight at a time.  The positions of each halo
// cell's first POS_SLOTS particles are read once into LDS, so that a group's
// record loads wait for one global read, not two.
constexpr int GATHER_THREADS = 512;
constexpr int HALO_CELLS = DEP_HALO * DEP_HALO;
static_assert(HALO_CELLS <= GATHER_THREADS, "one halo cell per thread");
static_assert(GATHER_THREADS == 2 * DEP_TILE * DEP_TILE, "two threads per output cell");
constexpr int COST_CLASSES = SMALL_CELL + 2;   // 0 .. SMALL_CELL particles, more
constexpr int POS_SLOTS = 8;
template <typename T> struct GatherPlan;
template <> struct GatherPlan<float> {
    static constexpr int g0 = 3, g1 = 2, g2 = 0, g3 = 0, g4 = 0, rf = 4, waves = 4;
    static constexpr bool rt = false;
};
template <> struct GatherPlan<double> {
    static constexpr int g0 = 1, g1 = 1, g2 = 1, g3 = 1, g4 = 1, rf = 4, waves = 4;
    static constexpr bool rt = true;   // one x offset at a time, in a runtime loop
};

// Binning position of particle k of a cell's list: from the cell's LDS row
// (spos) for its first POS_SLOTS particles, else from pos.
__device__ __forceinline__ int list_pos(const int *spos, const int *__restrict__ pos, int first,
                                        int k)
{
    return k < POS_SLOTS ? spos[k] : pos[first + k];
}

// Tap A (0..4) of weights4 for a wave-uniform A (a scalar branch).
template <typename T> __device__ __forceinline__ T w4_uniform(int A, T x)
{
    switch (A) {
    case 0: return w4_m2(x);
    case 1: return w4_m1(x);
    case 2: return w4_0(x);
    case 3: return w4_p1(x);
    default: return w4_p2(x);
    }
}

// cell_taps for the single, wave-uniform x offset A.
template <typename T>
__device__ __forceinline__ void cell_taps_uniform(int A, const DepRec<T> *__restrict__ rec,
                                                  const int *spos, const int *__restrict__ pos,
                                                  const T *__restrict__ dense, int first, int n,
                                                  int dsl, T inv_h, Tap4<T> (&tp)[5])
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
    constexpr int RF = GatherPlan<T>::rf;
    for (int j0 = 0; j0 < n; j0 += RF) {
        DepRec<T> rk[RF];
#pragma unroll
        for (int k = 0; k < RF; ++k)
            if (j0 + k < n) rk[k] = rec[list_pos(spos, pos, first, j0 + k)];
#pragma unroll
        for (int k = 0; k < RF; ++k) {
            if (j0 + k >= n) break;
            const DepRec<T> &d = rk[k];
            // The weight product first, then the currents (plasma3d/weights.py:58-67).
            T wy[5];
            weights4<T>(d.yl, wy);
            const T wx = w4_uniform<T>(A, d.xl) * inv_h;
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                const T w = wx * (wy[b] * inv_h);
                tp[b].v[0] += d.c0 * w;
                tp[b].v[1] += d.c1 * w;
                tp[b].v[2] += d.c2 * w;
                tp[b].v[3] += d.c3 * w;
            }
        }
    }
}

// Adds one particle's taps toward x offsets A0 .. A0 + G - 1 (target cells
// (ix + A - 2, iy + b - 2)): tp[A - A0][b].v[q] += c_q (w_A(xl) / h)
// (w_b(yl) / h).
template <int A0, int G, typename T>
__device__ __forceinline__ void add_taps(const DepRec<T> &d, T inv_h, Tap4<T> (&tp)[G][5])
{
    T wx[5], wy[5];
    weights4<T>(d.xl, wx);   // the taps outside the group are dropped by the compiler
    weights4<T>(d.yl, wy);
#pragma unroll
    for (int b = 0; b < 5; ++b) wy[b] *= inv_h;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const T w = wx[A0 + g] * inv_h;
        const T c0 = d.c0 * w, c1 = d.c1 * w, c2 = d.c2 * w, c3 = d.c3 * w;
#pragma unroll
        for (int b = 0; b < 5; ++b) {
            tp[g][b].v[0] += c0 * wy[b];
            tp[g][b].v[1] += c1 * wy[b];
            tp[g][b].v[2] += c2 * wy[b];
            tp[g][b].v[3] += c3 * wy[b];
        }
    }
}

// The taps of one source cell toward x offsets A0 .. A0 + G - 1: the sum
// over its particles, in list order, of the records the list names (pos,
// first = the cell's list start), or the precomputed block of a dense cell
// (dsl >= 0).
template <int A0, int G, typename T>
__device__ __forceinline__ void cell_taps(const DepRec<T> *__restrict__ rec, const int *spos,
                                          const int *__restrict__ pos,
                                          const T *__restrict__ dense, int first, int n, int dsl,
                                          T inv_h, Tap4<T> (&tp)[G][5])
{
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int b = 0; b < 5; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) tp[g][b].v[q] = T(0);
    if (dsl >= 0) {
        const T *v = dense + (size_t)dsl * DENSE_VALUES + 4 * (5 * A0);
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int b = 0; b < 5; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) tp[g][b].v[q] = v[20 * g + 4 * b + q];
        return;
    }
    // Up to RF records in flight at a time.
    constexpr int RF = GatherPlan<T>::rf;
    for (int j0 = 0; j0 < n; j0 += RF) {
        DepRec<T> d[RF];
#pragma unroll
        for (int k = 0; k < RF; ++k)
            if (j0 + k < n) d[k] = rec[list_pos(spos, pos, first, j0 + k)];
#pragma unroll
        for (int k = 0; k < RF; ++k) {
            if (j0 + k >= n) break;
            add_taps<A0, G, T>(d[k], inv_h, tp);
        }
    }
}

template <typename T> struct alignas(2 * sizeof(T)) Tap2 {
    T v[2];
};

// Every output cell is the fixed-order sum (x offset A ascending, then y
// offset b ascending) of the 25 taps that reach it, added in float64 by the
// two threads that own its four currents (thread t: cell t % 256, currents
// 2 (t / 256) and 2 (t / 256) + 1), and written by them alone: no
// floating-point atomics, reproducible bit for bit.  The LDS tap tile always
// spans the full 20 x 20 halo; its cells outside the grid hold zero taps, so
// that the reads need no bounds checks.
template <typename TP, typename TG>
__global__ __launch_bounds__(GATHER_THREADS, GatherPlan<TP>::waves) void k_deposit_gather(
    int N, double h, const int *__restrict__ start, BinState bs, TG *__restrict__ ro,
    TG *__restrict__ jx, TG *__restrict__ jy, TG *__restrict__ jz, const TG *__restrict__ bg_ro,
    TG *__restrict__ ro1, TG *__restrict__ jx1, TG *__restrict__ jy1, TG *__restrict__ jz1)
{
    using T = TP;   // arithmetic type of the taps
    using P = GatherPlan<T>;
    static_assert(P::g0 + P::g1 + P::g2 + P::g3 + P::g4 == 5, "five x offsets");
    constexpr int H = DEP_HALO;
    // One x offset's taps per halo cell, currents (ro, jx) and (jy, jz) apart.
    __shared__ Tap2<T> s_tap[2][5][HALO_CELLS];
    // List start, and particle count n or, for a dense cell, -1 - its slot.
    __shared__ int s_first[HALO_CELLS], s_nd[HALO_CELLS];
    __shared__ int s_pos[HALO_CELLS][POS_SLOTS];   // binning positions of its first particles
    __shared__ short s_cell[HALO_CELLS];       // halo cells by particle count
    __shared__ int s_class[COST_CLASSES];

    // The list sort has read the scan's tile bases: clear them for the next scan.
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < bs.nbase; i += GATHER_THREADS) bs.base[i] = 0;

    const int tile = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int X0 = (tile / tpr) * DEP_TILE, Y0 = (tile % tpr) * DEP_TILE;
    const DepRec<T> *__restrict__ rec = static_cast<const DepRec<T> *>(bs.rec);
    const T *__restrict__ dense = static_cast<const T *>(bs.dense);
    const int t = threadIdx.x;

    // Halo cell t (row t / 20, column t % 20 from (X0 - 2, Y0 - 2)): list
    // start, particle count, dense slot (or -1); none for a cell outside the
    // grid.  Then the cells ordered by cost class (particle count, a dense
    // cell counting as one) by a counting sort in LDS.  The order within a
    // class is arrival order, which decides only which thread forms which
    // cell's taps.
    if (t < COST_CLASSES) s_class[t] = 0;
    __syncthreads();
    int cls = 0, in_cls = 0;
    if (t < HALO_CELLS) {
        const int cx = X0 - 2 + t / H, cy = Y0 - 2 + t % H;
        int first = 0, n = 0, dsl = -1;
        if (cx >= 0 && cx < N && cy >= 0 && cy < N) {
            const int cell = cx * N + cy;
            first = start[cell];
            n = start[cell + 1] - first;
            dsl = bs.dslot[cell];
        }
        s_first[t] = first;
        s_nd[t] = dsl >= 0 ? -1 - dsl : n;
        if (dsl < 0) {
#pragma unroll
            for (int k = 0; k < POS_SLOTS; ++k)
                if (k < n) s_pos[t][k] = bs.pos[first + k];
        }
        cls = dsl >= 0 ? 1 : min(n, SMALL_CELL + 1);
        in_cls = atomicAdd(&s_class[cls], 1);
    }
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int c = 0; c < COST_CLASSES; ++c) {
            const int k = s_class[c];
            s_class[c] = run;
            run += k;
        }
    }
    __syncthreads();
    if (t < HALO_CELLS) s_cell[s_class[cls] + in_cls] = (short)t;
    __syncthreads();
    const int mine = t < HALO_CELLS ? s_cell[t] : -1;
    const int m_first = mine >= 0 ? s_first[mine] : 0;
    const int m_nd = mine >= 0 ? s_nd[mine] : 0;
    const int m_n = m_nd >= 0 ? m_nd : 0;
    const int m_dsl = m_nd < 0 ? -1 - m_nd : -1;

    const int o = t % (DEP_TILE * DEP_TILE), pair = t / (DEP_TILE * DEP_TILE);
    const int tx = o / DEP_TILE, ty = o % DEP_TILE;
    const int ox = X0 + tx, oy = Y0 + ty;
    // Source cell (ox - (A - 2), oy - (b - 2)) sits at halo position
    // base + (4 - A) H + (4 - b).
    const int base = tx * H + ty;
    const T inv_h = T(1.0 / h);
    double acc0 = 0, acc1 = 0;   // currents 2 pair and 2 pair + 1

    auto group = [&](auto a0_const, auto g_const) {
        constexpr int A0 = decltype(a0_const)::value, G = decltype(g_const)::value;
        Tap4<T> tp[G][5];
        if (mine >= 0)
            cell_taps<A0, G, T>(rec, s_pos[mine], bs.pos, dense, m_first, m_n, m_dsl, inv_h, tp);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (mine >= 0) {
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    s_tap[0][b][mine] = Tap2<T>{{tp[g][b].v[0], tp[g][b].v[1]}};
                    s_tap[1][b][mine] = Tap2<T>{{tp[g][b].v[2], tp[g][b].v[3]}};
                }
            }
            __syncthreads();
            const Tap2<T> *row = &s_tap[pair][0][base + (4 - (A0 + g)) * H];
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                const Tap2<T> v = row[b * HALO_CELLS + 4 - b];
                acc0 += (double)v.v[0];
                acc1 += (double)v.v[1];
            }
            __syncthreads();
        }
    };
    if constexpr (P::rt) {
#pragma unroll 1
        for (int A = 0; A < 5; ++A) {
            Tap4<T> tp[5];
            if (mine >= 0)
                cell_taps_uniform<T>(A, rec, s_pos[mine], bs.pos, dense, m_first, m_n, m_dsl,
                                     inv_h, tp);
            if (mine >= 0) {
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    s_tap[0][b][mine] = Tap2<T>{{tp[b].v[0], tp[b].v[1]}};
                    s_tap[1][b][mine] = Tap2<T>{{tp[b].v[2], tp[b].v[3]}};
                }
            }
            __syncthreads();
            const Tap2<T> *row = &s_tap[pair][0][base + (4 - A) * H];
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                const Tap2<T> v = row[b * HALO_CELLS + 4 - b];
                acc0 += (double)v.v[0];
                acc1 += (double)v.v[1];
            }
            __syncthreads();
        }
    } else {
    group(std::integral_constant<int, 0>{}, std::integral_constant<int, P::g0>{});
    group(std::integral_constant<int, P::g0>{}, std::integral_constant<int, P::g1>{});
    if constexpr (P::g2 > 0)
        group(std::integral_constant<int, P::g0 + P::g1>{}, std::integral_constant<int, P::g2>{});
    if constexpr (P::g3 > 0)
        group(std::integral_constant<int, P::g0 + P::g1 + P::g2>{},
              std::integral_constant<int, P::g3>{});
    if constexpr (P::g4 > 0)
        group(std::integral_constant<int, P::g0 + P::g1 + P::g2 + P::g3>{},
              std::integral_constant<int, P::g4>{});
    }

    if (ox < N && oy < N) {
        const int oc = ox * N + oy;
        if (pair == 0) {
            ro[oc] = (TG)acc0;
            jx[oc] = (TG)acc1;
        } else {
            jy[oc] = (TG)acc0;
            jz[oc] = (TG)acc1;
        }
        if (bg_ro != nullptr) {
            if (pair == 0) {
                ro1[oc] = bg_ro[oc];
                jx1[oc] = 0;
            } else {
                jy1[oc] = 0;
                jz1[oc] = 0;
            }
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
    const int ntiles = scan_tiles(ncell);   // cells 0..ncell
    if (ntiles != bs.nbase) return hipErrorInvalidValue;
    hipError_t e;
    if (timer && (e = timer->begin(KC_DEPOSIT, 1, s)) != hipSuccess) return e;
    if (g.n > 0) {
        if (!binned)
            k_bin<TP><<<pblocks, 256, 0, s>>>(g.n, g.h, N, a.x_init, a.y_init, xo, yo, px, py, pz,
                                              q, a.m, bs);
        k_scan_cells<<<ntiles, SCAN_THREADS, 0, s>>>(ncell, ntiles, bs.count, bs.lstart,
                                                     bs.base);
        k_scatter<<<pblocks, 256, 0, s>>>(g.n, ncell, bs);
        k_order_cells<<<(ncell + 255) / 256, 256, 0, s>>>(ncell, bs);
        // One block per CU (a block takes 134 KB of LDS): the queue is striding.
        k_dense_cells<TP><<<256, DENSE_THREADS, 0, s>>>(g.h, bs);
    } else {
        // Empty sort: every cell list is empty (and no cell is dense).
        k_scan_cells<<<ntiles, SCAN_THREADS, 0, s>>>(ncell, ntiles, bs.count, bs.lstart,
                                                     bs.base);
        k_order_cells<<<(ncell + 255) / 256, 256, 0, s>>>(ncell, bs);
    }
    const int tpr = (N + DEP_TILE - 1) / DEP_TILE;
    const int tiles = tpr * tpr;
    k_deposit_gather<TP, TG><<<tiles, GATHER_THREADS, 0, s>>>(N, g.h, bs.start, bs, ro, jx, jy,
                                                              jz, bg_ro, ro1, jx1, jy1, jz1);
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
