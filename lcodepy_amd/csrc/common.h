// Shared device helpers for the lcodepy_amd HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "kernels.h"

namespace lc3d {

// Sticky error bits the kernels raise; the host reads them back on demand.
enum : unsigned {
    ERR_PARTICLE_OFF_GRID = 1u,   // a plasma particle's 5x5 stencil left the grid
    ERR_NONFINITE = 2u,           // a non-finite plasma particle position
    ERR_BEAM_OFF_GRID = 4u,       // a beam particle's stencil left the grid (or
                                  // its substep count exceeded the cap)
    ERR_CELL_OVERFLOW = 8u,       // > 16384 plasma particles in one cell: that
                                  // cell's list is summed in arrival order
    ERR_BEAM_RANGE = 16u,         // a beam deposit contribution of 2^40 or more in
                                  // magnitude (not deposited)
};

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using cplx_t = typename cplx<T>::type;

// Quartic B-spline weights for the five taps -2..2 around the particle's
// cell, x in [-1/2, 1/2) the offset from the cell centre
// (plasma3d/weights.py:12-31).  Division by 4 is exact; 1/6 and 1/24 are
// applied as multiplications.
template <typename T> __device__ __forceinline__ T w4_m2(T x) {
    T t = T(0.5) - x, t2 = t * t;
    return t2 * t2 * T(1.0 / 24.0);
}
template <typename T> __device__ __forceinline__ T w4_m1(T x) {
    T x2 = x * x;
    return T(19.0 / 96.0) - T(11.0 / 24.0) * x + x2 * T(0.25) + x * x2 * T(1.0 / 6.0)
           - x2 * x2 * T(1.0 / 6.0);
}
template <typename T> __device__ __forceinline__ T w4_0(T x) {
    T x2 = x * x;
    return T(115.0 / 192.0) - x2 * T(0.625) + x2 * x2 * T(0.25);
}
template <typename T> __device__ __forceinline__ T w4_p1(T x) {
    T x2 = x * x;
    return T(19.0 / 96.0) + T(11.0 / 24.0) * x + x2 * T(0.25) - x * x2 * T(1.0 / 6.0)
           - x2 * x2 * T(1.0 / 6.0);
}
template <typename T> __device__ __forceinline__ T w4_p2(T x) {
    T t = T(0.5) + x, t2 = t * t;
    return t2 * t2 * T(1.0 / 24.0);
}
// Tap I (0..4, offset I - 2) of weights4.
template <int I, typename T> __device__ __forceinline__ T w4_tap(T x) {
    if constexpr (I == 0) return w4_m2(x);
    else if constexpr (I == 1) return w4_m1(x);
    else if constexpr (I == 2) return w4_0(x);
    else if constexpr (I == 3) return w4_p1(x);
    else return w4_p2(x);
}

template <typename T> __device__ __forceinline__ void weights4(T x, T w[5]) {
    w[0] = w4_m2(x);
    w[1] = w4_m1(x);
    w[2] = w4_0(x);
    w[3] = w4_p1(x);
    w[4] = w4_p2(x);
}

// Deposit-stencil coordinates of a plasma position (plasma3d/weights.py:
// 53-56 with x_h from weights.py:271-274): x_h = x / h + 1/2, cell
// floor(x_h) + N/2, offset x_h - floor(x_h) - 1/2.  Binning and deposit both
// call this, so a particle is always deposited around the cell it was binned
// into.  Returns false when the position is not finite or the 5x5 stencil
// would leave the grid.
__device__ __forceinline__ bool stencil_cell(double x, double y, double h, int N, int &ix,
                                             int &iy, double &xloc, double &yloc,
                                             bool &finite)
{
    const double xh = x / h + 0.5, yh = y / h + 0.5;
    finite = isfinite(xh) && isfinite(yh);
    if (!finite) return false;
    const double fx = floor(xh), fy = floor(yh);
    const double half = N / 2;
    if (fx + half < 2 || fx + half > N - 3 || fy + half < 2 || fy + half > N - 3) return false;
    ix = (int)(fx + half);
    iy = (int)(fy + half);
    xloc = xh - fx - 0.5;
    yloc = yh - fy - 0.5;
    return true;
}

// Maps a linear block id so that consecutive logical tiles land on the same
// XCD (the hardware deals workgroups round-robin over the 8 XCDs, each with
// its own L2).  n is the grid size; ids beyond the largest multiple of 8 keep
// their value.
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
    const int n8 = n & ~7;
    if (b >= n8) return b;
    return (b & 7) * (n8 >> 3) + (b >> 3);
}

// Counts the particles of a wave into their cells and returns each one's
// rank in its cell (arrival order): consecutive lanes with the same key (the
// common case when particles are visited in cell order) add their number
// with one atomic, whose old value their run head shares with them.  Every
// lane of the wave calls it; key < 0 is not counted and gets -1.
__device__ __forceinline__ int count_rank(int *count, int key)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const bool head = key >= 0 && (lane == 0 || prev != key);
    const unsigned long long ends = __ballot(head || key < 0);   // lanes that start a run
    int old = 0;
    if (head) {
        const unsigned long long above = ends & ~((2ull << lane) - 1ull);
        const int next = above ? __builtin_ctzll(above) : 64;
        old = atomicAdd(&count[key], next - lane);
    }
    const unsigned long long upto = ends & ((2ull << lane) - 1ull);   // run starts <= lane
    const int h = 63 - __builtin_clzll(upto);                        // this lane's run head
    const int base = __shfl(old, h);
    return key >= 0 ? base + (lane - h) : -1;
}

// Bins a plasma particle for the deposit: returns its cell key ix N + iy and
// writes its record at rec[j] (its binning position), or returns -1 (with an error bit) when the 5x5
// stencil would leave the grid or the position is not finite.  P Q Z M qv are
// the stored px py pz m q.
template <typename TP>
__device__ __forceinline__ int bin_particle(double x, double y, double h, int N, double P, double Q,
                                            double Z, double M, double qv, DepRec<TP> *rec,
                                            int64_t j, unsigned *err)
{
    int ix, iy;
    double xl, yl;
    bool finite;
    if (!stencil_cell(x, y, h, N, ix, iy, xl, yl, finite)) {
        atomicOr(err, finite ? ERR_PARTICLE_OFF_GRID : ERR_NONFINITE);
        return -1;
    }
    const double gm = sqrt(M * M + P * P + Q * Q + Z * Z);
    double dro, r_g;
    if constexpr (std::is_same<TP, float>::value) {
        // float32 records: one division by gamma, as a reciprocal
        const double ig = 1 / gm;
        dro = qv / (1 - Z * ig);
        r_g = dro * ig;
    } else {
        dro = qv / (1 - Z / gm);
        r_g = dro / gm;
    }
    DepRec<TP> r;
    r.xl = (TP)xl;
    r.yl = (TP)yl;
    r.c0 = (TP)dro;
    r.c1 = (TP)(P * r_g);
    r.c2 = (TP)(Q * r_g);
    r.c3 = (TP)(Z * r_g);
    rec[j] = r;
    return ix * N + iy;
}

}  // namespace lc3d
