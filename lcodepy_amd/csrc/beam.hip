// Beam layer kernels (beam3d), gfx950.
//
// A beam layer is small (thousands of macro-particles per xi step against
// ~10^6 plasma particles): one particle per thread would leave most of the
// GPU idle and each thread a long chain of dependent field loads, so a group
// of BEAM_LANES lanes serves one particle, each lane taking every
// BEAM_LANES-th of the 25 stencil taps.  The deposit adds each tap exactly,
// as a 128-bit fixed-point integer (below), so its result does not depend on
// the order of the atomics; the push (the reference's CPU explicit-midpoint scheme with
// its substep loop) sums the lanes' tap contributions by a butterfly over
// the group, which leaves the same bits in every lane, and every lane then
// runs the same momentum update.
#include "common.h"
#include "kernels.h"

namespace lc3d {

constexpr int MAX_BEAM_SUBSTEPS = 1 << 16;
constexpr int BEAM_LANES = 8;   // lanes per beam particle (a power of two dividing 64)
constexpr int BEAM_THREADS = 256;

// ---- exact fixed-point accumulation ------------------------------------------------
// A beam deposit adds each tap's charge v as the integer nearest to v 2^80,
// split into three 42-bit limbs (bits 0-41, 42-83, 84 up, all carrying v's
// sign), each added to its own 64-bit integer sum by an atomic whose result
// is not waited for.  Integer sums do not depend on the order of the
// additions, so a layer's density is the same bits whatever order the
// atomics land in; a limb sum absorbs 2^21 full-size additions per cell
// before it could overflow.  k_beam_finalize normalises the carries, turns
// each cell's sum into one float64 and adds it to the layer's density.
// Resolution 2^-80 (8e-25) of the unscaled charge; a contribution of 2^40 or
// more is refused (ERR_BEAM_RANGE).
constexpr int BEAM_FIX = 80, BEAM_LIMB = 42;
constexpr double BEAM_FIX_LIMIT = 1099511627776.0;   // 2^40

__device__ __forceinline__ void fixed_add(unsigned long long *l0, unsigned long long *l1,
                                          unsigned long long *l2, double v)
{
    // Each step below is exact: power-of-two scalings, floors, and
    // differences of a value and a lower multiple of the limb it stays
    // within; the lowest limb is rounded to the nearest integer.
    const double a = fabs(ldexp(v, BEAM_FIX));
    const double h2 = floor(ldexp(a, -2 * BEAM_LIMB));
    const double r2 = a - ldexp(h2, 2 * BEAM_LIMB);
    const double h1 = floor(ldexp(r2, -BEAM_LIMB));
    const double h0 = rint(r2 - ldexp(h1, BEAM_LIMB));
    long long q0 = (long long)h0, q1 = (long long)h1, q2 = (long long)h2;
    if (v < 0) {
        q0 = -q0;
        q1 = -q1;
        q2 = -q2;
    }
    atomicAdd(l0, (unsigned long long)q0);
    atomicAdd(l1, (unsigned long long)q1);
    atomicAdd(l2, (unsigned long long)q2);
}

// The float64 nearest (up to two roundings) to the limb sums' value times
// 2^-80: carries moved up so that the low limbs lie in [0, 2^42), then the
// three exact partial values added from the top.
__device__ __forceinline__ double fixed_value(unsigned long long l0, unsigned long long l1,
                                              unsigned long long l2)
{
    long long a0 = (long long)l0, a1 = (long long)l1, a2 = (long long)l2;
    const long long c0 = a0 >> BEAM_LIMB;   // arithmetic shift: floor division
    a0 -= c0 << BEAM_LIMB;
    a1 += c0;
    const long long c1 = a1 >> BEAM_LIMB;
    a1 -= c1 << BEAM_LIMB;
    a2 += c1;
    const double hi = ldexp((double)a2, 2 * BEAM_LIMB);
    const double mid = ldexp((double)a1, BEAM_LIMB);
    return ldexp((hi + mid) + (double)a0, -BEAM_FIX);
}

// beam3d/weights.py:40-64 with x_h and xi_loc from :152-160.  Layer 0 is the
// running layer (weight xi_loc), layer 1 the next one (weight 1 - xi_loc);
// acc holds the fixed-point limb sums, word w of cell o at acc[w N*N + o]
// (w = 0-2 the running layer's limbs, 3-5 the next layer's).
__global__ __launch_bounds__(BEAM_THREADS) void k_beam_deposit(BeamGeom g, int64_t n,
                                                               double xi_layer,
                                                               const double *__restrict__ x,
                                                               const double *__restrict__ y,
                                                               const double *__restrict__ xi,
                                                               const double *__restrict__ q_norm,
                                                               unsigned long long *__restrict__ acc,
                                                               unsigned *err)
{
    const int64_t t = blockIdx.x * (int64_t)BEAM_THREADS + threadIdx.x;
    const int64_t k = t / BEAM_LANES;
    const int lane = (int)(t % BEAM_LANES);
    if (k >= n) return;
    const int N = g.N;
    const int64_t nn = (int64_t)N * N;
    const double xh = x[k] / g.h + 0.5, yh = y[k] / g.h + 0.5;
    if (!isfinite(xh) || !isfinite(yh)) {
        if (lane == 0) atomicOr(err, ERR_BEAM_OFF_GRID);
        return;
    }
    const double fx = floor(xh), fy = floor(yh);
    const int ix = (int)fmin(fmax(fx + N / 2, -8.0), (double)N + 8);
    const int iy = (int)fmin(fmax(fy + N / 2, -8.0), (double)N + 8);
    const double xi_loc = (xi_layer - xi[k]) / g.xi_step;
    double wx[5], wy[5];
    weights4<double>(xh - fx - 0.5, wx);
    weights4<double>(yh - fy - 0.5, wy);
    const double q = q_norm[k];
    bool off = false, range = false;
    for (int tap = lane; tap < 25; tap += BEAM_LANES) {
        const int kx = tap / 5, ky = tap - 5 * kx;
        const int gx = ix + kx - 2, gy = iy + ky - 2;
        if (gx < 0 || gx >= N || gy < 0 || gy >= N) {
            off = true;
            continue;
        }
        const double w = wx[kx] * wy[ky];
        const double v0 = q * (w * xi_loc), v1 = q * (w * (1 - xi_loc));
        if (!(fabs(v0) < BEAM_FIX_LIMIT && fabs(v1) < BEAM_FIX_LIMIT)) {
            range = true;
            continue;
        }
        const int64_t o = (int64_t)gx * N + gy;
        fixed_add(&acc[o], &acc[nn + o], &acc[2 * nn + o], v0);
        fixed_add(&acc[3 * nn + o], &acc[4 * nn + o], &acc[5 * nn + o], v1);
    }
    if (off) atomicOr(err, ERR_BEAM_OFF_GRID);
    if (range) atomicOr(err, ERR_BEAM_RANGE);
}

// Adds each cell's fixed-point sums to ro0 / ro1 and clears them for the
// next deposit.  LAYOUT (beam_calculator.py:82-93): ro0 is the finished
// layer, (ro0 + sum0) / h2, and ro1 = sum1, at every cell.
template <bool LAYOUT>
__global__ __launch_bounds__(256) void k_beam_finalize(int64_t nn, double h2,
                                                       unsigned long long *__restrict__ acc,
                                                       double *__restrict__ ro0,
                                                       double *__restrict__ ro1)
{
    const int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (o >= nn) return;
    double sum[2] = {0.0, 0.0};
    bool any[2] = {false, false};
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        unsigned long long *w = acc + 3 * l * nn + o;
        const unsigned long long a0 = w[0], a1 = w[nn], a2 = w[2 * nn];
        if (a0 | a1 | a2) {
            any[l] = true;
            sum[l] = fixed_value(a0, a1, a2);
            w[0] = 0;
            w[nn] = 0;
            w[2 * nn] = 0;
        }
    }
    if (LAYOUT) {
        ro0[o] = (any[0] ? ro0[o] + sum[0] : ro0[o]) / h2;
        ro1[o] = sum[1];
    } else {
        if (any[0]) ro0[o] += sum[0];
        if (any[1]) ro1[o] += sum[1];
    }
}

// beam3d/beam_calculator.py:18-34, 97-123: dt multiplier halved until
// below sqrt(sqrt(1/q_m^2 + pz^2) / E_sub); dt = mult * time_step and
// remaining_steps = int(1 / mult).
__global__ __launch_bounds__(256) void k_beam_substep(int64_t n, const double *__restrict__ q_m,
                                                      const double *__restrict__ pz, double e_sub,
                                                      double time_step, double *__restrict__ dt,
                                                      int64_t *__restrict__ steps)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double max_dt = sqrt(sqrt(1 / (q_m[k] * q_m[k]) + pz[k] * pz[k]) / e_sub);
    double d = 1.0;
    // At most 2^-40: the push below also caps its substep loop.
    for (int it = 0; it < 40 && d > max_dt; ++it) d /= 2.0;
    dt[k] = d * time_step;
    steps[k] = (int64_t)(1. / d);
}

// This lane's share of the two-layer field gather of beam3d/move.py:70-91
// (taps lane, lane + BEAM_LANES, ...), summed over the particle's lanes.
template <typename TF>
__device__ __forceinline__ void gather2(const Fields6<TF> &a, const Fields6<TF> &b, int N, int ix,
                                        int iy, const double wx[5], const double wy[5], double w0s,
                                        double w1s, int lane, double F[6], bool &off)
{
    for (int tap = lane; tap < 25; tap += BEAM_LANES) {
        const int kx = tap / 5, ky = tap - 5 * kx;
        const int gx = ix + kx - 2, gy = iy + ky - 2;
        if (gx < 0 || gx >= N || gy < 0 || gy >= N) {
            off = true;
            continue;
        }
        const int o = gx * N + gy;
        const double w = wx[kx] * wy[ky];
        const double w0 = w * w0s, w1 = w * w1s;
        F[0] += (double)a.Ex[o] * w0 + (double)b.Ex[o] * w1;
        F[1] += (double)a.Ey[o] * w0 + (double)b.Ey[o] * w1;
        F[2] += (double)a.Ez[o] * w0 + (double)b.Ez[o] * w1;
        F[3] += (double)a.Bx[o] * w0 + (double)b.Bx[o] * w1;
        F[4] += (double)a.By[o] * w0 + (double)b.By[o] * w1;
        F[5] += (double)a.Bz[o] * w0 + (double)b.Bz[o] * w1;
    }
#pragma unroll
    for (int off_l = 1; off_l < BEAM_LANES; off_l <<= 1) {
#pragma unroll
        for (int c = 0; c < 6; ++c) F[c] += __shfl_xor(F[c], off_l, BEAM_LANES);
    }
}

// beam3d/move.py:31-127 (the CPU scheme; its trailing "fell" test compares
// x, not xi, as move.py:123 does).
// A group of BEAM_LANES lanes per particle; all lanes of a group run the
// same substep loop (the same particle state), so the group's shuffles see
// every lane of it; lane 0 writes the particle back.
template <typename TF>
__global__ __launch_bounds__(BEAM_THREADS) void k_beam_move(
    BeamGeom g, int64_t n, double xi_k, double xi_k1, const double *__restrict__ q_m,
    const double *__restrict__ dt, Fields6<TF> fk, Fields6<TF> fk1, int64_t *__restrict__ steps,
    int64_t *__restrict__ id, double *__restrict__ x, double *__restrict__ y,
    double *__restrict__ xi, double *__restrict__ px, double *__restrict__ py,
    double *__restrict__ pz, uint8_t *__restrict__ lost, uint8_t *__restrict__ moved,
    uint8_t *__restrict__ fell, unsigned *err)
{
    const int64_t t = blockIdx.x * (int64_t)BEAM_THREADS + threadIdx.x;
    const int64_t k = t / BEAM_LANES;
    const int lane = (int)(t % BEAM_LANES);
    if (k >= n) return;   // whole groups: n particles fill BEAM_LANES n lanes
    const int N = g.N;
    const double r2 = g.lost_radius * g.lost_radius;
    const double qm = q_m[k], iqm = 1 / qm, d = dt[k];
    const double sgn = qm < 0 ? -1.0 : (qm > 0 ? 1.0 : 0.0);
    double X = x[k], Y = y[k], XI = xi[k], PX = px[k], PY = py[k], PZ = pz[k];
    int64_t rs = steps[k];
    bool is_lost = false, is_fell = false, off = false;
    // Substep counts come from the caller; cap the loop so that a corrupt
    // count cannot keep the wave running (flagged as an error).
    int guard = 0;
    while (rs > 0) {
        if (++guard > MAX_BEAM_SUBSTEPS) {
            off = true;
            break;
        }
        double gm = sqrt(iqm * iqm + PX * PX + PY * PY + PZ * PZ);
        const double xh_ = X + d / 2 * (PX / gm);
        const double yh_ = Y + d / 2 * (PY / gm);
        const double xih = XI + d / 2 * (PZ / gm - 1);
        if (xi_k1 > xih) {
            is_fell = true;
            break;
        }
        if (xh_ * xh_ + yh_ * yh_ >= r2) {
            X = xh_;
            Y = yh_;
            XI = xih;
            is_lost = true;
            rs = 0;
            break;
        }
        const double xhh = xh_ / g.h + .5, yhh = yh_ / g.h + .5;
        const double fx = floor(xhh), fy = floor(yhh);
        const int ix = (int)fmin(fmax(fx + N / 2, -8.0), (double)N + 8);
        const int iy = (int)fmin(fmax(fy + N / 2, -8.0), (double)N + 8);
        double wx[5], wy[5];
        weights4<double>(xhh - fx - .5, wx);
        weights4<double>(yhh - fy - .5, wy);
        const double xi_loc = (xih - xi_k) / g.xi_step;
        double F[6] = {0, 0, 0, 0, 0, 0};
        gather2<TF>(fk, fk1, N, ix, iy, wx, wy, xi_loc, 1 - xi_loc, lane, F, off);
        const double vx = PX / gm, vy = PY / gm, vz = PZ / gm;
        const double a = sgn * d / 2;
        const double pxh = PX + a * (F[0] + vy * F[5] - vz * F[4]);
        const double pyh = PY + a * (F[1] + vz * F[3] - vx * F[5]);
        const double pzh = PZ + a * (F[2] + vx * F[4] - vy * F[3]);
        gm = sqrt(iqm * iqm + pxh * pxh + pyh * pyh + pzh * pzh);
        X += d * (pxh / gm);
        Y += d * (pyh / gm);
        XI += d * (pzh / gm - 1);
        PX = 2 * pxh - PX;
        PY = 2 * pyh - PY;
        PZ = 2 * pzh - PZ;
        if (X * X + Y * Y >= r2) {
            is_lost = true;
            rs = 0;
            break;
        }
        rs -= 1;
    }
    if (off) atomicOr(err, ERR_BEAM_OFF_GRID);
    if (lane != 0) return;
    if (rs == 0 && xi_k1 > X) is_fell = true;
    if (is_lost) id[k] = -id[k];
    x[k] = X;
    y[k] = Y;
    xi[k] = XI;
    px[k] = PX;
    py[k] = PY;
    pz[k] = PZ;
    steps[k] = rs;
    lost[k] = is_lost;
    fell[k] = is_fell;
    moved[k] = !is_fell && !is_lost;
}

// beam_calculator.py:124-139 (the split of a moved layer): for each group g
// (lost, moved, fell; blockIdx.y) the particles whose flag[g] is set are
// copied, in layer order, to the front of out_f[g] / out_i[g], and their
// number is written to counts[g].  A block takes a chunk of SPLIT_THREADS
// particles and first counts the selected particles before its chunk (at
// most SPLIT_MAX_BLOCKS chunks, so a few KB of flags), or, launched with one
// block per group, walks all chunks; within a chunk a wave's selected lanes
// are ranked by a ballot, the waves' totals by one thread.
constexpr int SPLIT_THREADS = 1024;
constexpr int SPLIT_MAX_BLOCKS = 32;
__global__ __launch_bounds__(SPLIT_THREADS) void k_beam_split(int64_t n, int nf, int ni,
                                                              const double *__restrict__ f,
                                                              const int64_t *__restrict__ iv,
                                                              const uint8_t *__restrict__ flags,
                                                              double *__restrict__ out_f,
                                                              int64_t *__restrict__ out_i,
                                                              int64_t *__restrict__ counts)
{
    __shared__ int wave_off[SPLIT_THREADS / 64 + 1];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint8_t *fl = flags + (int64_t)g * n;
    double *of = out_f + (int64_t)g * nf * n;
    int64_t *oi = out_i + (int64_t)g * ni * n;
    const bool one = gridDim.x == 1;
    const int64_t c_begin = one ? 0 : (int64_t)blockIdx.x * SPLIT_THREADS;
    const int64_t c_end = one ? n : min(n, c_begin + SPLIT_THREADS);
    int64_t base = 0;
    if (!one) {
        // Selected particles before this chunk: per wave by ballots, then
        // the waves' sums by one thread.
        int cnt = 0;
        for (int64_t k0 = 0; k0 < c_begin; k0 += SPLIT_THREADS)
            cnt += __builtin_popcountll(__ballot(fl[k0 + threadIdx.x] != 0));
        if (lane == 0) wave_off[wave] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int w = 0; w < SPLIT_THREADS / 64; ++w) acc += wave_off[w];
            wave_off[SPLIT_THREADS / 64] = acc;
        }
        __syncthreads();
        base = wave_off[SPLIT_THREADS / 64];
        __syncthreads();
    }
    for (int64_t c0 = c_begin; c0 < c_end; c0 += SPLIT_THREADS) {
        const int64_t k = c0 + threadIdx.x;
        const bool sel = k < c_end && fl[k] != 0;
        const unsigned long long m = __ballot(sel);
        const int before = __builtin_popcountll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_off[wave] = __builtin_popcountll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int w = 0; w < SPLIT_THREADS / 64; ++w) {
                const int t = wave_off[w];
                wave_off[w] = acc;
                acc += t;
            }
            wave_off[SPLIT_THREADS / 64] = acc;
        }
        __syncthreads();
        if (sel) {
            const int64_t d = base + wave_off[wave] + before;
            for (int r = 0; r < nf; ++r) of[(int64_t)r * n + d] = f[(int64_t)r * n + k];
            for (int r = 0; r < ni; ++r) oi[(int64_t)r * n + d] = iv[(int64_t)r * n + k];
        }
        base += wave_off[SPLIT_THREADS / 64];
        __syncthreads();
    }
    if (threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) counts[g] = base;
}

hipError_t launch_beam_split(int64_t n, int nf, int ni, const double *f, const int64_t *iv,
                             const uint8_t *flags, double *out_f, int64_t *out_i, int64_t *counts,
                             hipStream_t s)
{
    const int64_t chunks = (n + SPLIT_THREADS - 1) / SPLIT_THREADS;
    const unsigned bx = chunks >= 2 && chunks <= SPLIT_MAX_BLOCKS ? (unsigned)chunks : 1u;
    k_beam_split<<<dim3(bx, 3), SPLIT_THREADS, 0, s>>>(n, nf, ni, f, iv, flags, out_f, out_i,
                                                        counts);
    return hipGetLastError();
}

hipError_t launch_beam_deposit(const BeamGeom &g, int64_t n, int64_t plasma_layer_idx,
                               const double *x, const double *y, const double *xi,
                               const double *q_norm, unsigned long long *acc, double *ro0,
                               double *ro1, unsigned *err, hipStream_t s, bool layout)
{
    if (n <= 0 && !layout) return hipSuccess;
    const double xi_layer = -g.xi_step * (double)plasma_layer_idx;
    if (n > 0) {
        const int64_t blocks = (n * BEAM_LANES + BEAM_THREADS - 1) / BEAM_THREADS;
        k_beam_deposit<<<(unsigned)blocks, BEAM_THREADS, 0, s>>>(g, n, xi_layer, x, y, xi, q_norm,
                                                                 acc, err);
    }
    const int64_t nn = (int64_t)g.N * g.N;
    const unsigned fb = (unsigned)((nn + 255) / 256);
    if (layout) k_beam_finalize<true><<<fb, 256, 0, s>>>(nn, g.h * g.h, acc, ro0, ro1);
    else k_beam_finalize<false><<<fb, 256, 0, s>>>(nn, g.h * g.h, acc, ro0, ro1);
    return hipGetLastError();
}

hipError_t launch_beam_substep(int64_t n, const double *q_m, const double *pz,
                               double substepping_energy, double time_step, double *dt,
                               int64_t *remaining_steps, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_beam_substep<<<(int)((n + 255) / 256), 256, 0, s>>>(n, q_m, pz, substepping_energy,
                                                           time_step, dt, remaining_steps);
    return hipGetLastError();
}

template <typename TF>
hipError_t launch_beam_move(const BeamGeom &g, int64_t n, int64_t beam_layer_idx,
                            const double *q_m, const double *dt, const Fields6<TF> &fk,
                            const Fields6<TF> &fk1, int64_t *remaining_steps, int64_t *id,
                            double *x, double *y, double *xi, double *px, double *py, double *pz,
                            uint8_t *lost, uint8_t *moved, uint8_t *fell, unsigned *err,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const double xi_k = (double)beam_layer_idx * -g.xi_step;
    const double xi_k1 = (double)(beam_layer_idx + 1) * -g.xi_step;
    const int64_t blocks = (n * BEAM_LANES + BEAM_THREADS - 1) / BEAM_THREADS;
    k_beam_move<TF><<<(unsigned)blocks, BEAM_THREADS, 0, s>>>(g, n, xi_k, xi_k1, q_m, dt, fk, fk1,
                                                              remaining_steps, id, x, y, xi, px,
                                                              py, pz, lost, moved, fell, err);
    return hipGetLastError();
}

template hipError_t launch_beam_move<double>(const BeamGeom &, int64_t, int64_t, const double *,
                                             const double *, const Fields6<double> &,
                                             const Fields6<double> &, int64_t *, int64_t *,
                                             double *, double *, double *, double *, double *,
                                             double *, uint8_t *, uint8_t *, uint8_t *,
                                             unsigned *, hipStream_t);

}  // namespace lc3d
