// Plasma particle push for one xi step (gfx950).
//
// One thread per macro-particle over structure-of-arrays buffers, so every
// load and store of the particle state is a fully coalesced 64-lane access.
// The push follows plasma3d/move.py:52-148: midpoint position (reflected at
// the wall), 5x5 quartic-spline gather of six fields, two fixed-point passes
// of the relativistic momentum update, position update from the half-step
// momentum, wall reflection and absorption.  In the FP32-bulk form the
// momentum arithmetic runs in float while gamma, the 1/(1 - v_z) factor and
// the positions stay double.
#include "common.h"
#include "kernels.h"

#include <type_traits>

namespace lc3d {

template <typename TP>
__global__ __launch_bounds__(256) void k_move_wo_fields(int64_t n, double xi_step,
                                                        const TP *__restrict__ m,
                                                        double *__restrict__ xo,
                                                        double *__restrict__ yo,
                                                        const TP *__restrict__ px,
                                                        const TP *__restrict__ py,
                                                        const double *__restrict__ pz)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double mk = m[k];
    if (mk == 0) return;
    const double a = px[k], b = py[k], c = pz[k];
    const double g = sqrt(mk * mk + c * c + a * a + b * b);
    const double d = g - c;
    xo[k] += a / d * xi_step;
    yo[k] += b / d * xi_step;
}

// Reflection of a position at +-b (plasma3d/move.py:28-49).  Returns true
// when any reflection happened (the caller then zeroes the momenta).
__device__ __forceinline__ bool reflect(double x_init, double y_init, double b, double &x,
                                        double &y, double &xo, double &yo)
{
    bool hit = false;
    if (x > b) { x = 2 * b - x; xo = x - x_init; hit = true; }
    if (x < -b) { x = -2 * b - x; xo = x - x_init; hit = true; }
    if (y > b) { y = 2 * b - y; yo = y - y_init; hit = true; }
    if (y < -b) { y = -2 * b - y; yo = y - y_init; hit = true; }
    return hit;
}

// Pushes particle k; with BIN returns its deposit cell (or -1) after writing
// its deposit record and cell at binning position j, otherwise -1.
template <typename TP, typename TF, bool FIRST, bool BIN>
__device__ __forceinline__ int push_one(const PushGeom &g, const Fields6<TF> &f,
                                        const SortArrays<TP> &a, const BinState &bs, int64_t k,
                                        int64_t j)
{
    const int N = g.N;
    const double b = g.boundary, dxi = g.xi_step;
    // FP32-bulk (float32 momenta): quotients that share a divisor are taken
    // as products with its reciprocal (one float64 division instead of two
    // or three; the results differ by at most an ulp of float64, far below
    // the float32 momenta).  The float64 form divides as
    // plasma3d/move.py:52-148 does.
    constexpr bool RCP = std::is_same<TP, float>::value;

    const double xi0 = a.x_init[k], yi0 = a.y_init[k];
    const double xop = a.xo_prev[k], yop = a.yo_prev[k];
    const TP opx = a.px_prev[k], opy = a.py_prev[k];
    const double opz = a.pz_prev[k];
    const double m = (double)a.m[k];

    TP q;
    double xof, yof;
    if (FIRST) {
        // particles_full = copy(particles_prev) + free-streaming estimate
        q = a.q_prev[k];
        xof = xop;
        yof = yop;
        if (m != 0) {
            const double A = opx, B = opy, C = opz;
            const double gm = sqrt(m * m + C * C + A * A + B * B);
            if constexpr (RCP) {
                const double r = dxi / (gm - C);
                xof = xop + A * r;
                yof = yop + B * r;
            } else {
                xof = xop + A / (gm - C) * dxi;
                yof = yop + B / (gm - C) * dxi;
            }
        }
    } else {
        q = a.q[k];
        xof = a.xo[k];
        yof = a.yo[k];
    }

    if (q == TP(0)) {
        if (FIRST) {
            a.xo[k] = xof;
            a.yo[k] = yof;
            a.px[k] = opx;
            a.py[k] = opy;
            a.pz[k] = opz;
            a.q[k] = q;
        }
        if (BIN) bs.key[j] = -1;
        return -1;
    }

    // Midpoint between the previous position and the current estimate.
    double x = xi0 + (xop + xof) * 0.5, y = yi0 + (yop + yof) * 0.5;
    double xo_tmp = 0, yo_tmp = 0;
    TP px = opx, py = opy;
    double pz = opz;
    if (reflect(xi0, yi0, b, x, y, xo_tmp, yo_tmp)) {
        px = py = TP(0);
        pz = 0;
    }

    // 5x5 gather of the six fields at the midpoint (move.py:90-107).
    TP Ex = 0, Ey = 0, Ez = 0, Bx = 0, By = 0, Bz = 0;
    {
        int ix, iy;
        double xl, yl;
        bool finite;
        if (stencil_cell(x, y, g.h, N, ix, iy, xl, yl, finite)) {
            TP wx[5], wy[5];
            weights4<TP>((TP)xl, wx);
            weights4<TP>((TP)yl, wy);
#pragma unroll
            for (int kx = 0; kx < 5; ++kx) {
                const int row = (ix + kx - 2) * N + iy - 2;
#pragma unroll
                for (int ky = 0; ky < 5; ++ky) {
                    const TP w = wx[kx] * wy[ky];
                    const int id = row + ky;
                    Ex += (TP)f.Ex[id] * w;
                    Bx += (TP)f.Bx[id] * w;
                    Ey += (TP)f.Ey[id] * w;
                    By += (TP)f.By[id] * w;
                    Ez += (TP)f.Ez[id] * w;
                    Bz += (TP)f.Bz[id] * w;
                }
            }
        } else {
            // Outside the grid (or not finite): no field is gathered.
            atomicOr(bs.err, finite ? ERR_PARTICLE_OFF_GRID : ERR_NONFINITE);
        }
    }

    // Two fixed-point passes of the momentum update (move.py:110-125).
    const double qd = (double)q * dxi;
    TP dpx, dpy, dpz;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        const double P = px, Q = py, R = pz;
        const double gm = sqrt(m * m + R * R + P * P + Q * Q);
        double ux, uy, uz;
        if constexpr (RCP) {
            const double ig = 1 / gm;
            ux = P * ig;
            uy = Q * ig;
            uz = R * ig;
        } else {
            ux = P / gm;
            uy = Q / gm;
            uz = R / gm;
        }
        const TP vx = (TP)ux, vy = (TP)uy, vz = (TP)uz;
        const TP factor = (TP)(qd / (1 - uz));
        dpx = factor * (Ex + vy * Bz - vz * By);
        dpy = factor * (Ey - vx * Bz + vz * Bx);
        dpz = factor * (Ez + vx * By - vy * Bx);
        px = opx + dpx * TP(0.5);
        py = opy + dpy * TP(0.5);
        pz = opz + (double)(dpz * TP(0.5));
    }

    // Position from the half-step momentum, then the full-step momentum.
    double xo, yo;
    {
        const double P = px, Q = py, R = pz;
        const double gm = sqrt(m * m + R * R + P * P + Q * Q);
        const double d = gm - R;
        if constexpr (RCP) {
            const double r = dxi / d;
            xo = xop + P * r;
            yo = yop + Q * r;
        } else {
            xo = xop + P / d * dxi;
            yo = yop + Q / d * dxi;
        }
    }
    px = opx + dpx;
    py = opy + dpy;
    pz = opz + (double)dpz;

    x = xi0 + xo;
    y = yi0 + yo;
    if (reflect(xi0, yi0, b, x, y, xo, yo)) {
        px = py = TP(0);
        pz = 0;
    }
    x = xi0 + xo;
    y = yi0 + yo;
    bool absorbed = false;
    if (x >= b || x <= -b || y >= b || y <= -b) {
        xo = yo = 0;
        px = py = TP(0);
        pz = 0;
        absorbed = true;
    }

    a.xo[k] = xo;
    a.yo[k] = yo;
    a.px[k] = px;
    a.py[k] = py;
    a.pz[k] = pz;
    if (FIRST || absorbed) a.q[k] = absorbed ? TP(0) : q;

    int key = -1;
    if (BIN) {
        if (!absorbed)
            key = bin_particle<TP>(xi0 + xo, yi0 + yo, g.h, N, (double)px, (double)py,
                                   (double)pz, m, (double)q, static_cast<DepRec<TP> *>(bs.rec),
                                   j, bs.err);
        bs.key[j] = key;
    }
    return key;
}

template <typename TP, typename TF, bool FIRST, bool BIN>
__global__ __launch_bounds__(256) void k_push(PushGeom g, Fields6<TF> f, SortArrays<TP> a,
                                              BinState bs, const int *__restrict__ order)
{
    // Threads past the end stay to the count so that whole waves reach it.
    // Consecutive blocks of the order go to the same XCD, so each XCD's L2
    // fetches the field rows of its own stretch of cells.
    const int64_t j = xcd_swizzle(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x;
    int key = -1;
    int64_t k = 0;
    if (j < g.n) {
        k = order ? (int64_t)order[j] : j;   // the particle this thread pushes
        key = push_one<TP, TF, FIRST, BIN>(g, f, a, bs, k, j);
        if (BIN) bs.pid[j] = (int)k;
    }
    if (BIN) {
        // Unbinned particles are counted (and ranked) in count[N * N].
        const int ck = j < g.n ? (key >= 0 ? key : g.N * g.N) : -1;
        const int r = count_rank(bs.count, ck);
        if (j < g.n) bs.rank[j] = r;
    }
}



template <typename TP>
hipError_t launch_move_wo_fields(int64_t n, double xi_step, const TP *m, double *xo, double *yo,
                                 const TP *px, const TP *py, const double *pz, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (int)((n + 255) / 256);
    k_move_wo_fields<TP><<<blocks, 256, 0, s>>>(n, xi_step, m, xo, yo, px, py, pz);
    return hipGetLastError();
}

template <typename TP, typename TF>
hipError_t launch_push(const PushGeom &g, const Fields6<TF> &f, const SortArrays<TP> &a,
                       bool first, bool bin, const BinState &bs, const int *order, hipStream_t s,
                       KernelTimer *timer)
{
    if (g.n <= 0) return hipSuccess;
    const int blocks = (int)((g.n + 255) / 256);
    hipError_t e;
    if (timer && (e = timer->begin(KC_PUSH, 1, s)) != hipSuccess) return e;
    if (first && bin) k_push<TP, TF, true, true><<<blocks, 256, 0, s>>>(g, f, a, bs, order);
    else if (first) k_push<TP, TF, true, false><<<blocks, 256, 0, s>>>(g, f, a, bs, order);
    else if (bin) k_push<TP, TF, false, true><<<blocks, 256, 0, s>>>(g, f, a, bs, order);
    else k_push<TP, TF, false, false><<<blocks, 256, 0, s>>>(g, f, a, bs, order);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timer) return timer->end(s);
    return hipSuccess;
}

// Rounds the six float64 fields into float32 copies, field f at f32 + f
// stride (stride: N*N rounded up to a multiple of 4, so that every row of
// four starts 16-byte aligned); four elements per thread.
__global__ __launch_bounds__(256) void k_fields_to_f32(int64_t nn, int64_t stride,
                                                       Fields6<double> f, float *__restrict__ out)
{
    const int64_t o = 4 * (blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
    if (o >= nn) return;
    const double *src[6] = {f.Ex, f.Ey, f.Ez, f.Bx, f.By, f.Bz};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float *d = out + k * stride + o;
        if (o + 4 <= nn) {
            const double *p = src[k] + o;
            const float4 v = make_float4((float)p[0], (float)p[1], (float)p[2], (float)p[3]);
            *reinterpret_cast<float4 *>(d) = v;
        } else {
            for (int64_t e = o; e < nn; ++e) out[k * stride + e] = (float)src[k][e];
        }
    }
}

hipError_t launch_push_f64_fields_as_f32(const PushGeom &g, const Fields6<double> &f,
                                         float *f32, const SortArrays<float> &a, bool first,
                                         bool bin, const BinState &bs, const int *order,
                                         hipStream_t s, KernelTimer *timer)
{
    if (g.n <= 0) return hipSuccess;
    const int64_t nn = (int64_t)g.N * g.N, st = fields32_stride(g.N);
    const int blocks = (int)((g.n + 255) / 256);
    const Fields6<float> f2{f32, f32 + st, f32 + 2 * st, f32 + 3 * st, f32 + 4 * st, f32 + 5 * st};
    hipError_t e;
    if (timer && (e = timer->begin(KC_PUSH, 1, s)) != hipSuccess) return e;
    k_fields_to_f32<<<(unsigned)((nn / 4 + 256) / 256), 256, 0, s>>>(nn, st, f, f32);
    if (first && bin) k_push<float, float, true, true><<<blocks, 256, 0, s>>>(g, f2, a, bs, order);
    else if (first) k_push<float, float, true, false><<<blocks, 256, 0, s>>>(g, f2, a, bs, order);
    else if (bin) k_push<float, float, false, true><<<blocks, 256, 0, s>>>(g, f2, a, bs, order);
    else k_push<float, float, false, false><<<blocks, 256, 0, s>>>(g, f2, a, bs, order);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timer) return timer->end(s);
    return hipSuccess;
}

template hipError_t launch_move_wo_fields<double>(int64_t, double, const double *, double *,
                                                  double *, const double *, const double *,
                                                  const double *, hipStream_t);
template hipError_t launch_move_wo_fields<float>(int64_t, double, const float *, double *,
                                                 double *, const float *, const float *,
                                                 const double *, hipStream_t);
template hipError_t launch_push<double, double>(const PushGeom &, const Fields6<double> &,
                                                const SortArrays<double> &, bool, bool,
                                                const BinState &, const int *, hipStream_t, KernelTimer *);
template hipError_t launch_push<float, double>(const PushGeom &, const Fields6<double> &,
                                               const SortArrays<float> &, bool, bool,
                                               const BinState &, const int *, hipStream_t, KernelTimer *);
template hipError_t launch_push<float, float>(const PushGeom &, const Fields6<float> &,
                                              const SortArrays<float> &, bool, bool,
                                              const BinState &, const int *, hipStream_t, KernelTimer *);

}  // namespace lc3d
