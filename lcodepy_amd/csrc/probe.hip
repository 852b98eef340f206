// f(xi) probe diagnostics on the device (diagnostics/FXi/Strategy3D.py:28-95),
// gfx950.
//
// The reference reads grid values at its probe points (and the maximum of
// |dx_chaotic| ...) into host lists after every xi step.  Here one launch per
// xi step samples every requested source grid at every probe point and
// writes the values as float64 into a row of a device buffer the caller
// keeps for the whole time step, so the xi loop never copies to the host;
// the row is read back once, when the diagnostic is dumped.  The maxima of
// |a| over particle arrays use a fixed-order two-stage reduction (max is
// order-independent anyway; the fixed order keeps NaN handling the same
// from run to run).
#include "common.h"
#include "kernels.h"

namespace lc3d {

// One thread per (source, probe point).  Out-of-grid probe points read NaN
// (the host rejects them before upload; this keeps the kernel in bounds
// whatever the index arrays hold).
__global__ __launch_bounds__(256) void k_probe_sample(ProbeSources ps, int N, int n_points,
                                                      const int *__restrict__ ix,
                                                      const int *__restrict__ iy,
                                                      double *__restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ps.n * n_points) return;
    const int s = t / n_points, p = t % n_points;
    const int x = ix[p], y = iy[p];
    double v = __builtin_nan("");
    if (x >= 0 && x < N && y >= 0 && y < N) {
        const size_t o = (size_t)x * N + y;
        v = ps.is_f32[s] ? (double)static_cast<const float *>(ps.src[s])[o]
                         : static_cast<const double *>(ps.src[s])[o];
    }
    out[t] = v;
}

constexpr int PROBE_THREADS = 256;

__device__ __forceinline__ double block_max(double v, double *red)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int off = PROBE_THREADS / 2; off > 0; off >>= 1) {
        if (t < off) red[t] = fmax(red[t], red[t + off]);
        __syncthreads();
    }
    const double m = red[0];
    __syncthreads();
    return m;
}

// max |a[k]| over the rows [r0, r1) x columns [c0, c1) of an (nr, nc)
// row-major array (the whole array, or the central window the reference's
// *_chaotic_perp quantities take, Strategy3D.py:89-95).
__global__ __launch_bounds__(PROBE_THREADS) void k_absmax_partial(const double *__restrict__ a,
                                                                  int64_t nc, int64_t r0,
                                                                  int64_t r1, int64_t c0,
                                                                  int64_t c1,
                                                                  double *__restrict__ part)
{
    __shared__ double red[PROBE_THREADS];
    const int64_t w = c1 - c0, n = (r1 - r0) * w;
    double m = 0;
    for (int64_t k = blockIdx.x * (int64_t)PROBE_THREADS + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * PROBE_THREADS) {
        const int64_t r = r0 + k / w, c = c0 + k % w;
        m = fmax(m, fabs(a[r * nc + c]));
    }
    m = block_max(m, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(PROBE_THREADS) void k_absmax_final(int nb,
                                                                const double *__restrict__ part,
                                                                double *__restrict__ out)
{
    __shared__ double red[PROBE_THREADS];
    double m = 0;
    for (int b = threadIdx.x; b < nb; b += PROBE_THREADS) m = fmax(m, part[b]);
    m = block_max(m, red);
    if (threadIdx.x == 0) out[0] = m;
}

hipError_t launch_probe_sample(const ProbeSources &ps, int N, int n_points, const int *ix,
                               const int *iy, double *out, hipStream_t s)
{
    const int total = ps.n * n_points;
    if (total <= 0) return hipSuccess;
    k_probe_sample<<<(total + 255) / 256, 256, 0, s>>>(ps, N, n_points, ix, iy, out);
    return hipGetLastError();
}

hipError_t launch_absmax(const double *a, int64_t nc, int64_t r0, int64_t r1, int64_t c0,
                         int64_t c1, double *part, int max_blocks, double *out, hipStream_t s)
{
    const int64_t n = std::max<int64_t>(0, (r1 - r0) * (c1 - c0));
    const int nb = (int)std::max<int64_t>(
        1, std::min<int64_t>(max_blocks, (n + PROBE_THREADS - 1) / PROBE_THREADS));
    k_absmax_partial<<<nb, PROBE_THREADS, 0, s>>>(a, nc, r0, r1, c0, c1, part);
    k_absmax_final<<<1, PROBE_THREADS, 0, s>>>(nb, part, out);
    return hipGetLastError();
}

}  // namespace lc3d
