// Declustering noise filter of the plasma electrons (plasma3d/noise_filter.py:
// 53-238), gfx950.  The sort's arrays are its (n, n) lattice in row-major
// order, element (i, j) at i n + j.
//
//   k_dc_inhom  inhomogeneity of x_offt and y_offt along both lattice
//               directions (noise_filter.py:53-76, 135-136);
//   k_dc_rows   circular Gaussian smoothing along j;
//   k_dc_cols   the same along i, and the chaotic displacement
//               dx_chaotic = inhomogeneity - smoothed (noise_filter.py:147-170);
//   k_dc_kick   restoring kicks of px and py from the forces of the particle
//               and its four lattice neighbours, and the damping
//               (noise_filter.py:176-232).
//
// The reference smooths with ifft2(fourier_gaussian(fft2(a), sigma)).real:
// multiplying the periodic spectrum by exp(-2 pi^2 sigma^2 (k/n)^2) along
// each axis.  That is a circular convolution with the kernel
// g[m] = (1/n) sum_k exp(-2 pi^2 sigma^2 (k/n)^2) cos(2 pi k m / n), applied
// along j and then along i; the host computes g in float64 and keeps the
// taps above 1e-17 of g[0] (the rest change no float64 result).  Every
// output is one thread's fixed-order sum.  The kick arithmetic is evaluated
// operation by operation as the reference's numpy statements (no fused
// multiply-adds), in float64; px, py are read and written in the bulk type.
#include "common.h"
#include "kernels.h"

namespace lc3d {

constexpr int DC_THREADS = 256;

// Inhomogeneity of d at (i, j) along j (`along_j`) or along i
// (calculate_inhomogeneity_numba with axis 0 or 1).
__device__ __forceinline__ double dc_inhom(const double *__restrict__ d, int n, int i, int j,
                                           bool along_j)
{
    const int64_t o = (int64_t)i * n + j;
    const int64_t st = along_j ? 1 : n;
    const int k = along_j ? j : i;
    if (k >= 1 && k <= n - 2) {
        const double c = d[o], l = d[o - st], r = d[o + st];
        return (c != 0 && l != 0 && r != 0) ? c - (l + r) / 2 : 0.0;
    }
    const int64_t s = k == 0 ? st : -st;   // towards the inside
    const double a = d[o], b = d[o + s], c = d[o + 2 * s];
    return (a != 0 && b != 0 && c != 0) ? a - 2 * b + c : 0.0;
}

__global__ __launch_bounds__(DC_THREADS) void k_dc_inhom(int n, const double *__restrict__ xo,
                                                         const double *__restrict__ yo,
                                                         double *__restrict__ dxi,
                                                         double *__restrict__ dyi)
{
    const int64_t o = blockIdx.x * (int64_t)DC_THREADS + threadIdx.x;
    if (o >= (int64_t)n * n) return;
    const int i = (int)(o / n), j = (int)(o % n);
    dxi[o] = (dc_inhom(xo, n, i, j, true) + dc_inhom(xo, n, i, j, false)) / 2;
    dyi[o] = (dc_inhom(yo, n, i, j, false) + dc_inhom(yo, n, i, j, true)) / 2;
}

// Circular convolution along j of both inputs, one block per lattice row
// (staged in LDS): out[i][j] = sum_t taps[t] in[i][(j - tap0 - t) mod n].
constexpr int DC_MAX_SIDE = 4096;
__global__ __launch_bounds__(DC_THREADS) void k_dc_rows(int n, int ntaps, int tap0,
                                                        const double *__restrict__ taps,
                                                        const double *__restrict__ ax,
                                                        const double *__restrict__ ay,
                                                        double *__restrict__ tx,
                                                        double *__restrict__ ty)
{
    __shared__ double rx[DC_MAX_SIDE], ry[DC_MAX_SIDE];
    const int64_t row = (int64_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += DC_THREADS) {
        rx[j] = ax[row + j];
        ry[j] = ay[row + j];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += DC_THREADS) {
        int s = (j - tap0) % n;
        if (s < 0) s += n;
        double sx = 0, sy = 0;
        for (int t = 0; t < ntaps; ++t) {
            const double g = taps[t];
            sx += g * rx[s];
            sy += g * ry[s];
            s = s == 0 ? n - 1 : s - 1;
        }
        tx[row + j] = sx;
        ty[row + j] = sy;
    }
}

// The same along i, then dx_chaotic = inhomogeneity - smoothed.
__global__ __launch_bounds__(DC_THREADS) void k_dc_cols(int n, int ntaps, int tap0,
                                                        const double *__restrict__ taps,
                                                        const double *__restrict__ tx,
                                                        const double *__restrict__ ty,
                                                        const double *__restrict__ dxi,
                                                        const double *__restrict__ dyi,
                                                        double *__restrict__ dxc,
                                                        double *__restrict__ dyc)
{
    const int64_t o = blockIdx.x * (int64_t)DC_THREADS + threadIdx.x;
    if (o >= (int64_t)n * n) return;
    const int i = (int)(o / n), j = (int)(o % n);
    int s = (i - tap0) % n;
    if (s < 0) s += n;
    double sx = 0, sy = 0;
    for (int t = 0; t < ntaps; ++t) {
        const double g = taps[t];
        const int64_t q = (int64_t)s * n + j;
        sx += g * tx[q];
        sy += g * ty[q];
        s = s == 0 ? n - 1 : s - 1;
    }
    dxc[o] = dxi[o] - sx;
    dyc[o] = dyi[o] - sy;
}

struct KickCoef {
    double xi;        // xi_step
    double mforce;    // -declustering-force
    double xidamp;    // xi_step * damping
    double lim2;      // declustering-limit ** 2
};

// d_limit and the force of the particle at o (noise_filter.py:170-179).
__device__ __forceinline__ void dc_force(const KickCoef &k, double cx, double cy, double &fx,
                                         double &fy, double &dl)
{
#pragma clang fp contract(off)
    const double d = cx * cx + cy * cy;
    dl = d < k.lim2 ? 1 - d / k.lim2 : 0.0;
    fx = k.mforce * cx * dl;
    fy = k.mforce * cy * dl;
}

// px of one particle after noise_filter.py:218-224 (likewise py): the seven
// in-place array statements applied to it in their order.
__device__ __forceinline__ double dc_kick(const KickCoef &k, double p, int i, int j, int n,
                                          double f, double f_im, double f_ip, double f_jm,
                                          double f_jp, double dl, double c, double cprev)
{
#pragma clang fp contract(off)
    if (i >= 1 && i <= n - 2) p += k.xi * f / 2;
    if (i >= 2) p -= k.xi * f_im / 4;
    if (i <= n - 3) p -= k.xi * f_ip / 4;
    if (j >= 1 && j <= n - 2) p += k.xi * f / 2;
    if (j >= 2) p -= k.xi * f_jm / 4;
    if (j <= n - 3) p -= k.xi * f_jp / 4;
    p -= k.xidamp * dl * (c - cprev);
    return p;
}

template <typename TP>
__global__ __launch_bounds__(DC_THREADS) void k_dc_kick(int n, KickCoef k,
                                                        const double *__restrict__ dxc,
                                                        const double *__restrict__ dyc,
                                                        const double *__restrict__ dxp,
                                                        const double *__restrict__ dyp,
                                                        TP *__restrict__ px, TP *__restrict__ py)
{
    const int64_t o = blockIdx.x * (int64_t)DC_THREADS + threadIdx.x;
    if (o >= (int64_t)n * n) return;
    const int i = (int)(o / n), j = (int)(o % n);
    double fx, fy, dl;
    dc_force(k, dxc[o], dyc[o], fx, fy, dl);
    // Forces of the four neighbours (0 where a neighbour is off the lattice;
    // those terms are not applied).
    double nx[4] = {0, 0, 0, 0}, ny[4] = {0, 0, 0, 0};
    const int64_t nb[4] = {o - n, o + n, o - 1, o + 1};
    const bool ok[4] = {i >= 1, i <= n - 2, j >= 1, j <= n - 2};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (ok[q]) {
            double d;
            dc_force(k, dxc[nb[q]], dyc[nb[q]], nx[q], ny[q], d);
        }
    }
    px[o] = (TP)dc_kick(k, (double)px[o], i, j, n, fx, nx[0], nx[1], nx[2], nx[3], dl, dxc[o], dxp[o]);
    py[o] = (TP)dc_kick(k, (double)py[o], i, j, n, fy, ny[0], ny[1], ny[2], ny[3], dl, dyc[o], dyp[o]);
}

template <typename TP>
hipError_t launch_noise_filter(const DeclustGeom &g, const double *xo, const double *yo, TP *px,
                               TP *py, const double *dxp, const double *dyp, double *dxc,
                               double *dyc, hipStream_t s)
{
    const int n = g.side;
    if (n < 3 || n > DC_MAX_SIDE || g.ntaps < 1) return hipErrorInvalidValue;
    const int64_t nn = (int64_t)n * n;
    const unsigned blocks = (unsigned)((nn + DC_THREADS - 1) / DC_THREADS);
    double *dxi = g.scratch, *dyi = g.scratch + nn, *tx = g.scratch + 2 * nn, *ty = g.scratch + 3 * nn;
    const KickCoef k{g.xi_step, -g.force, g.xi_step * g.damping, g.limit * g.limit};
    k_dc_inhom<<<blocks, DC_THREADS, 0, s>>>(n, xo, yo, dxi, dyi);
    k_dc_rows<<<n, DC_THREADS, 0, s>>>(n, g.ntaps, g.tap0, g.taps, dxi, dyi, tx, ty);
    k_dc_cols<<<blocks, DC_THREADS, 0, s>>>(n, g.ntaps, g.tap0, g.taps, tx, ty, dxi, dyi, dxc, dyc);
    k_dc_kick<TP><<<blocks, DC_THREADS, 0, s>>>(n, k, dxc, dyc, dxp, dyp, px, py);
    return hipGetLastError();
}

template hipError_t launch_noise_filter<double>(const DeclustGeom &, const double *, const double *,
                                                double *, double *, const double *, const double *,
                                                double *, double *, hipStream_t);
template hipError_t launch_noise_filter<float>(const DeclustGeom &, const double *, const double *,
                                               float *, float *, const double *, const double *,
                                               double *, double *, hipStream_t);

}  // namespace lc3d
