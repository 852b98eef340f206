// Diagnostics reductions on the device (gfx950).
//
// The total energy flux of a plasma layer, diagnostics/FXi/Strategy3D.py:6-25
// (calculate_energy_fluxes): the electromagnetic part
//     sum over the grid of (Ez^2 + Bz^2 + (Ex - By)^2 + (Ey + Bx)^2) h^2 / 2
// plus the electrons' motion energy sum of sqrt(m^2 + px^2 + py^2 + pz^2) - m.
// Two launches, no atomics: every block sums its strided share of grid points
// and particles into two float64 partials (a fixed-order tree in LDS), and
// one block adds the partials in a fixed order, so the result is
// reproducible bit for bit.
#include "common.h"
#include "kernels.h"

namespace lc3d {

constexpr int DIAG_THREADS = 256;

__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int off = DIAG_THREADS / 2; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

template <typename TP>
__global__ __launch_bounds__(DIAG_THREADS) void k_energy_partial(int N, double h, Fields6<double> f,
                                                                 int64_t n, const TP *__restrict__ m,
                                                                 const TP *__restrict__ px,
                                                                 const TP *__restrict__ py,
                                                                 const double *__restrict__ pz,
                                                                 double *__restrict__ part)
{
    __shared__ double red[DIAG_THREADS];
    const int64_t nn = (int64_t)N * N;
    const int64_t stride = (int64_t)gridDim.x * DIAG_THREADS;
    const int64_t t0 = blockIdx.x * (int64_t)DIAG_THREADS + threadIdx.x;
    const double w = h * h / 2;
    double em = 0, kin = 0;
    for (int64_t o = t0; o < nn; o += stride) {
        const double ez = f.Ez[o], bz = f.Bz[o];
        const double a = f.Ex[o] - f.By[o], b = f.Ey[o] + f.Bx[o];
        em += (ez * ez + bz * bz + a * a + b * b) * w;
    }
    for (int64_t k = t0; k < n; k += stride) {
        const double M = m[k], P = px[k], Q = py[k], R = pz[k];
        kin += sqrt(M * M + P * P + Q * Q + R * R) - M;
    }
    em = block_sum(em, red);
    kin = block_sum(kin, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = em;
        part[2 * blockIdx.x + 1] = kin;
    }
}

__global__ __launch_bounds__(DIAG_THREADS) void k_energy_final(int nb, const double *__restrict__ part,
                                                               double *__restrict__ psi)
{
    __shared__ double red[DIAG_THREADS];
    double em = 0, kin = 0;
    for (int b = threadIdx.x; b < nb; b += DIAG_THREADS) {
        em += part[2 * b];
        kin += part[2 * b + 1];
    }
    em = block_sum(em, red);
    kin = block_sum(kin, red);
    if (threadIdx.x == 0) psi[0] = kin + em;   // as Strategy3D.py:24
}

template <typename TP>
hipError_t launch_energy_flux(int N, double h, const Fields6<double> &f, int64_t n, const TP *m,
                              const TP *px, const TP *py, const double *pz, double *part,
                              int max_blocks, double *psi, hipStream_t s)
{
    const int64_t work = std::max<int64_t>((int64_t)N * N, n);
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(max_blocks, (work + DIAG_THREADS - 1) / DIAG_THREADS));
    k_energy_partial<TP><<<nb, DIAG_THREADS, 0, s>>>(N, h, f, n, m, px, py, pz, part);
    k_energy_final<<<1, DIAG_THREADS, 0, s>>>(nb, part, psi);
    return hipGetLastError();
}

template hipError_t launch_energy_flux<double>(int, double, const Fields6<double> &, int64_t,
                                               const double *, const double *, const double *,
                                               const double *, double *, int, double *, hipStream_t);
template hipError_t launch_energy_flux<float>(int, double, const Fields6<double> &, int64_t,
                                              const float *, const float *, const float *,
                                              const double *, double *, int, double *, hipStream_t);

}  // namespace lc3d
