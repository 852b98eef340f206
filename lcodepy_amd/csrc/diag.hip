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
#include <mutex>
#include <string>

#include "../../include/lcodepy_amd.h"
#include "common.h"
#include "kernels.h"

namespace lc3d {

constexpr int DIAG_THREADS = 256;

// Partial sums of the allocation-free energy flux (lc3d_energy_flux_noalloc):
// two float64 per block of the first launch, at most NOALLOC_BLOCKS blocks
// (16 KiB per device, held until the process ends).  Host code reaches it
// only through the address hipGetSymbolAddress returns on each device.
constexpr int NOALLOC_BLOCKS = LC3D_NOALLOC_FLUX_BLOCKS;
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS];

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

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[MAX_DEVICES] = {};
thread_local std::string g_noalloc_error = "null context";

int noalloc_fail(int code, const std::string &msg)
{
    g_noalloc_error = msg;
    return code;
}

// The device address of `scratch` on the current device, looked up on the
// first call there.
hipError_t scratch_on_device(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(lc3d::scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

const char *lc3d::noalloc_last_error() { return g_noalloc_error.c_str(); }

extern "C" int lc3d_energy_flux_noalloc(int32_t N, double h, int32_t precision,
                                        const lc3d_fields6 *f, int64_t n, const void *m,
                                        const void *px, const void *py, const double *pz,
                                        double *psi, void *stream)
{
    using namespace lc3d;
    if (N < 1 || N > LC3D_NOALLOC_FLUX_MAX_N || n < 0 || n > LC3D_NOALLOC_FLUX_MAX_PARTICLES)
        return noalloc_fail(LC3D_EINVAL, "energy_flux_noalloc: N must be in [1, " +
                                             std::to_string(LC3D_NOALLOC_FLUX_MAX_N) +
                                             "] and n in [0, " +
                                             std::to_string((long long)LC3D_NOALLOC_FLUX_MAX_PARTICLES) +
                                             "]");
    if ((precision != LC3D_FP64 && precision != LC3D_FP32_BULK) || !f || !f->Ex || !f->Ey ||
        !f->Ez || !f->Bx || !f->By || !f->Bz || !psi || (n > 0 && (!m || !px || !py || !pz)))
        return noalloc_fail(LC3D_EINVAL, "energy_flux_noalloc: bad arguments");
    double *part = nullptr;
    hipError_t e = scratch_on_device(&part);
    if (e != hipSuccess)
        return noalloc_fail(LC3D_EHIP, std::string("energy_flux_noalloc: scratch address: ") +
                                           hipGetErrorString(e));
    const Fields6<double> fl{static_cast<const double *>(f->Ex), static_cast<const double *>(f->Ey),
                             static_cast<const double *>(f->Ez), static_cast<const double *>(f->Bx),
                             static_cast<const double *>(f->By), static_cast<const double *>(f->Bz)};
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == LC3D_FP32_BULK)
        e = launch_energy_flux<float>(N, h, fl, n, static_cast<const float *>(m),
                                      static_cast<const float *>(px), static_cast<const float *>(py),
                                      pz, part, NOALLOC_BLOCKS, psi, st);
    else
        e = launch_energy_flux<double>(N, h, fl, n, static_cast<const double *>(m),
                                       static_cast<const double *>(px),
                                       static_cast<const double *>(py), pz, part, NOALLOC_BLOCKS,
                                       psi, st);
    if (e != hipSuccess)
        return noalloc_fail(LC3D_EHIP, std::string("energy_flux_noalloc: ") + hipGetErrorString(e));
    return LC3D_OK;
}
