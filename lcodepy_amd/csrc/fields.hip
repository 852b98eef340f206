// Field solve of one xi step (plasma3d/fields.py:112-322), gfx950.
//
// The seven fields are solved together, as three batched passes over
// N x N grids (N = 2M + 1... here M = N - 1 is the half period):
//
//   A  rows, forward:   right-hand side of every field computed on the fly
//                       from the currents (fields.py:112-266) and transformed
//                       along y (DST-I or DCT-I);
//   B  columns:         transform along x, multiply by 1/(l_i + l_j (+c))
//                       / (2M)^2, transform back along x -- all in LDS;
//   C  rows, inverse:   transform back along y and form the outputs
//                       (full = 2 half - prev for Ex Ey Bx By, half =
//                       (full + prev)/2 for Ez Bz, fields.py:299-314);
//   Bz finalize:        subtract the mean of Bz (fields.py:264).
//
// Each DST-I / DCT-I of length ~N is the real FFT of the odd / even
// extension of period 2M, computed as one complex FFT of length M (even and
// odd samples packed as real and imaginary parts) followed by the standard
// split.  The complex FFT is a self-sorting Stockham FFT in LDS with radix
// 2/4/8 stages plus generic 3/5/7/11/13 stages; one wavefront owns one
// transform.  Twiddles exp(-i pi t / M) are computed on the host in float64.
#include "solve_common.h"

namespace lc3d {

constexpr int WAVE = 64;
constexpr int WAVES = 4;   // transforms per workgroup
// Register arrays are sized by MAXM: the exact M for the lengths with a
// compile-time plan (StaticPlan below), otherwise the largest M a runtime-plan
// instance serves (64, 255, 511 or 1023).

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { C r; r.x = a.y; r.y = -a.x; return r; }  // a * (-i)

template <typename C> __device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3) {
    const C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <typename T, int R> struct Dft {
    // Generic odd radix: y_q = sum_p x_p w^(pq), w = exp(-2 pi i / R).
    using C = cplx_t<T>;
    static __device__ __forceinline__ void run(C (&v)[R], const C *__restrict__ tw, int M) {
        C y[R];
        const int step = 2 * (M / R);
#pragma unroll
        for (int qq = 0; qq < R; ++qq) {
            C s = v[0];
#pragma unroll
            for (int p = 1; p < R; ++p) s = cadd(s, cmul(v[p], tw[((p * qq) % R) * step]));
            y[qq] = s;
        }
#pragma unroll
        for (int qq = 0; qq < R; ++qq) v[qq] = y[qq];
    }
};
template <typename T> struct Dft<T, 2> {
    using C = cplx_t<T>;
    static __device__ __forceinline__ void run(C (&v)[2], const C *, int) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <typename T> struct Dft<T, 4> {
    using C = cplx_t<T>;
    static __device__ __forceinline__ void run(C (&v)[4], const C *, int) { dft4(v[0], v[1], v[2], v[3]); }
};
template <typename T> struct Dft<T, 8> {
    using C = cplx_t<T>;
    static __device__ __forceinline__ void run(C (&v)[8], const C *, int) {
        C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4(e0, e1, e2, e3);
        dft4(o0, o1, o2, o3);
        const T s = T(0.70710678118654752440);   // sqrt(1/2)
        C w1, w3;
        w1.x = s * (o1.x + o1.y);
        w1.y = s * (o1.y - o1.x);                // o1 * exp(-i pi/4)
        const C w2 = cmul_mi(o2);                // o2 * exp(-i pi/2)
        w3.x = s * (o3.y - o3.x);
        w3.y = -s * (o3.x + o3.y);               // o3 * exp(-3 i pi/4)
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, w1);
        v[5] = csub(e1, w1);
        v[2] = cadd(e2, w2);
        v[6] = csub(e2, w2);
        v[3] = cadd(e3, w3);
        v[7] = csub(e3, w3);
    }
};

// One Stockham stage of radix R on an M-point sequence in LDS, one wave.
template <typename T, int MAXM, int R>
__device__ __forceinline__ void fft_stage(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int M,
                                          int Ns, int lane)
{
    using C = cplx_t<T>;
    constexpr int MAXB = (MAXM / R + WAVE - 1) / WAVE;
    const int nb = M / R;
    C v[MAXB][R];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = lane + i * WAVE;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[b + r * nb];
        }
    }
    __syncthreads();
    const int stride = 2 * (M / (Ns * R));
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = lane + i * WAVE;
        if (b < nb) {
            const int k = b % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[r * k * stride]);
            Dft<T, R>::run(v[i], tw, M);
            const int base = (b / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns] = v[i][r];
        }
    }
    __syncthreads();
}

// Radix plans fixed at compile time for the transform lengths of the
// standard windows (N = 129, 257, 321, 385, 513, 769, 1025): every stage is
// unrolled with constant strides.  Other lengths use the runtime plan.
template <int M> struct StaticPlan { static constexpr int ns = 0; static constexpr int r[1] = {1}; };
template <> struct StaticPlan<128> { static constexpr int ns = 3; static constexpr int r[3] = {8, 8, 2}; };
template <> struct StaticPlan<256> { static constexpr int ns = 3; static constexpr int r[3] = {8, 8, 4}; };
template <> struct StaticPlan<320> { static constexpr int ns = 3; static constexpr int r[3] = {8, 8, 5}; };
template <> struct StaticPlan<384> { static constexpr int ns = 4; static constexpr int r[4] = {8, 8, 2, 3}; };
template <> struct StaticPlan<512> { static constexpr int ns = 3; static constexpr int r[3] = {8, 8, 8}; };
template <> struct StaticPlan<768> { static constexpr int ns = 4; static constexpr int r[4] = {8, 8, 4, 3}; };
template <> struct StaticPlan<1024> { static constexpr int ns = 4; static constexpr int r[4] = {8, 8, 4, 4}; };

template <typename T, int M, int R, int Ns>
__device__ __forceinline__ void fft_stage_c(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int lane)
{
    using C = cplx_t<T>;
    constexpr int nb = M / R;
    constexpr int MAXB = (nb + WAVE - 1) / WAVE;
    constexpr int stride = 2 * (M / (Ns * R));
    C v[MAXB][R];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = lane + i * WAVE;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[b + r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = lane + i * WAVE;
        if (b < nb) {
            const int k = b % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[r * k * stride]);
            Dft<T, R>::run(v[i], tw, M);
            const int base = (b / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int M, int S, int Ns>
__device__ __forceinline__ void fft_static(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int lane)
{
    if constexpr (S < StaticPlan<M>::ns) {
        constexpr int R = StaticPlan<M>::r[S];
        fft_stage_c<T, M, R, Ns>(buf, tw, lane);
        fft_static<T, M, S + 1, Ns * R>(buf, tw, lane);
    }
}

template <typename T, int MAXM>
__device__ void fft(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, const FftPlan &pl, int lane)
{
    if constexpr (StaticPlan<MAXM>::ns > 0) {
        fft_static<T, MAXM, 0, 1>(buf, tw, lane);   // MAXM is the exact length here
        return;
    }
    int Ns = 1;
    for (int s = 0; s < pl.ns; ++s) {
        const int R = pl.radix[s];
        switch (R) {
        case 2: fft_stage<T, MAXM, 2>(buf, tw, pl.M, Ns, lane); break;
        case 3: fft_stage<T, MAXM, 3>(buf, tw, pl.M, Ns, lane); break;
        case 4: fft_stage<T, MAXM, 4>(buf, tw, pl.M, Ns, lane); break;
        case 5: fft_stage<T, MAXM, 5>(buf, tw, pl.M, Ns, lane); break;
        case 7: fft_stage<T, MAXM, 7>(buf, tw, pl.M, Ns, lane); break;
        case 8: fft_stage<T, MAXM, 8>(buf, tw, pl.M, Ns, lane); break;
        case 11: fft_stage<T, MAXM, 11>(buf, tw, pl.M, Ns, lane); break;
        case 13: fft_stage<T, MAXM, 13>(buf, tw, pl.M, Ns, lane); break;
        default: break;
        }
        Ns *= R;
    }
}

// Unnormalised DST-I (x(1..M-1) -> Y(1..M-1)) or DCT-I (x(0..M) -> Y(0..M))
// of one sequence, computed by one wave.  load(t) gives x(t); store(k, Y)
// receives the result.  `active` false runs the barriers only.
template <typename T, int MAXM, typename Load, typename Store>
__device__ void r2r(bool dst, cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, const FftPlan &pl,
                    int lane, bool active, Load load, Store store)
{
    using C = cplx_t<T>;
    constexpr int MAXN = (MAXM + WAVE - 1) / WAVE;
    constexpr int MAXK = MAXM / WAVE + 1;
    const int M = StaticPlan<MAXM>::ns > 0 ? MAXM : pl.M;
    // Pack the period-2M extension: z[n] = e[2n] + i e[2n+1].
    C z[MAXN];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        const int n = lane + i * WAVE;
        if (active && n < M) {
            T e[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int tt = 2 * n + u;
                if (dst) {
                    e[u] = (tt == 0 || tt == M) ? T(0) : (tt < M ? load(tt) : -load(2 * M - tt));
                } else {
                    e[u] = tt <= M ? load(tt) : load(2 * M - tt);
                }
            }
            z[i].x = e[0];
            z[i].y = e[1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        const int n = lane + i * WAVE;
        if (active && n < M) buf[n] = z[i];
    }
    __syncthreads();
    fft<T, MAXM>(buf, tw, pl, lane);
    // Split: E[k] = (Z[k] + Z*[M-k])/2 + W^k (Z[k] - Z*[M-k])/(2i), W = exp(-i pi/M).
    T y[MAXK];
#pragma unroll
    for (int i = 0; i < MAXK; ++i) {
        const int k = lane + i * WAVE;
        y[i] = 0;
        if (active && k <= M) {
            const C zk = buf[k == M ? 0 : k];
            const C zm = buf[k == 0 ? 0 : M - k];
            const T ar = T(0.5) * (zk.x + zm.x), ai = T(0.5) * (zk.y - zm.y);
            const T br = T(0.5) * (zk.y + zm.y), bi = T(-0.5) * (zk.x - zm.x);
            const C w = tw[k];
            const T er = ar + (w.x * br - w.y * bi);
            const T ei = ai + (w.x * bi + w.y * br);
            y[i] = dst ? -ei : er;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXK; ++i) {
        const int k = lane + i * WAVE;
        if (active && k <= M && (!dst || (k > 0 && k < M))) store(k, y[i]);
    }
    __syncthreads();
}

// ---- pass A: right-hand sides + forward transform along y -------------------
__device__ __forceinline__ bool has_spectral_row(int f, int i, int N) {
    // Rows that carry data after the transform along y: all rows for the
    // fields whose x-transform is a DCT, the interior rows otherwise.
    return dst_along_x(f) ? (i > 0 && i < N - 1) : (i >= 0 && i < N);
}

template <typename TG, typename TPRED, int MAXM>
__global__ __launch_bounds__(256) void k_rows_fwd(SolveGeom g, SolveIn<TG, TPRED> in, int nrowblk)
{
    using C = cplx_t<TG>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    C *buf = reinterpret_cast<C *>(lds_raw) + wave * g.plan.M;
    const int f = blockIdx.y;
    const int i = blockIdx.x * WAVES + wave;
    const int N = g.N;
    const bool active = i < N && has_spectral_row(f, i, N);
    RhsEval<TG, TPRED> rhs{in, N, TG(1.0 / (2.0 * g.h)), TG(1.0 / g.xi_step), TG(g.c)};
    TG *out = reinterpret_cast<TG *>(g.work) + (size_t)f * N * N + (size_t)i * N;
    (void)nrowblk;
    r2r<TG, MAXM>(dst_along_y(f), buf, reinterpret_cast<const C *>(g.tw), g.plan, lane, active,
            [&](int t) { return rhs(f, i, t); }, [&](int k, TG y) { out[k] = y; });
}

// ---- pass B: columns, forward x-transform, multiply, inverse x-transform ------
template <typename TG, int MAXM>
__global__ __launch_bounds__(256) void k_cols(SolveGeom g)
{
    using C = cplx_t<TG>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    C *buf = reinterpret_cast<C *>(lds_raw) + wave * g.plan.M;
    const int f = blockIdx.y;
    const int N = g.N;
    const int j = xcd_swizzle(blockIdx.x, gridDim.x) * WAVES + wave;   // spectral column (ky)
    const bool ydst = dst_along_y(f), xdst = dst_along_x(f);
    const bool active = ydst ? (j > 0 && j < N - 1) : (j < N);
    TG *col = reinterpret_cast<TG *>(g.work) + (size_t)f * N * N + j;
    const double lj = active ? g.lam[j] : 0.0;
    const double cc = (f == F_EZ || f == F_BZ || f == F_PHI) ? 0.0 : g.c;
    const double d2 = 4.0 * (double)(N - 1) * (double)(N - 1);
    TG *bre = reinterpret_cast<TG *>(buf);
    // Forward along x, multiplied by the solver matrix (initialization.py:15-67)
    // on the way out; the spectrum is parked in the real view of the buffer.
    r2r<TG, MAXM>(xdst, buf, reinterpret_cast<const C *>(g.tw), g.plan, lane, active,
            [&](int t) { return col[(size_t)t * N]; },
            [&](int k, TG y) {
                const double den = g.lam[k] + lj + cc;
                const double mul = (f == F_BZ && k == 0 && j == 0) ? 0.0 : (1.0 / den) / d2;
                bre[k] = (TG)((double)y * mul);
            });
    r2r<TG, MAXM>(xdst, buf, reinterpret_cast<const C *>(g.tw), g.plan, lane, active,
            [&](int t) { return bre[t]; }, [&](int k, TG y) { col[(size_t)k * N] = y; });
}

// ---- pass C: inverse along y and output assembly --------------------------------
template <typename TG, typename TFULL, typename THALF, int MAXM>
__global__ __launch_bounds__(256) void k_rows_inv(SolveGeom g, SolveOut<TFULL, THALF> out,
                                                  bool write_full)
{
    using C = cplx_t<TG>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    C *buf = reinterpret_cast<C *>(lds_raw) + wave * g.plan.M;
    const int f = blockIdx.y;
    const int i = blockIdx.x * WAVES + wave;
    const int N = g.N;
    const bool row_ok = i < N;
    const bool spec = row_ok && has_spectral_row(f, i, N);
    const TG *src = reinterpret_cast<const TG *>(g.work) + (size_t)f * N * N + (size_t)i * N;
    const size_t ro = (size_t)i * N;

    auto emit = [&](int jj, TG r) { emit_output(out, f, ro + jj, r, write_full); };

    double s = 0;
    const bool ydst = dst_along_y(f);
    r2r<TG, MAXM>(ydst, buf, reinterpret_cast<const C *>(g.tw), g.plan, lane, spec,
            [&](int t) { return src[t]; },
            [&](int k, TG y) {
                emit(k, y);
                s += (double)y;
            });
    if (row_ok && !spec) {
        for (int jj = lane; jj < N; jj += WAVE) emit(jj, TG(0));   // zero rows of the solve
    } else if (spec && ydst && lane == 0) {
        emit(0, TG(0));
        emit(N - 1, TG(0));
    }
    if (f == F_BZ && row_ok) {
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) g.rowsum[i] = s;
    }
}

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields(const SolveGeom &g, const SolveIn<TG, TPRED> &in,
                                 const SolveOut<TFULL, THALF> &out, bool write_full, hipStream_t s)
{
    const int N = g.N;
    const size_t lds = (size_t)WAVES * g.plan.M * sizeof(cplx_t<TG>);
    const dim3 rows((N + WAVES - 1) / WAVES, 7);
    const int M = g.plan.M;
#define LC3D_SOLVE_STATIC(MM)                                                          \
    case MM:                                                                          \
        k_rows_fwd<TG, TPRED, MM><<<rows, 256, lds, s>>>(g, in, rows.x);              \
        k_cols<TG, MM><<<rows, 256, lds, s>>>(g);                                      \
        k_rows_inv<TG, TFULL, THALF, MM><<<rows, 256, lds, s>>>(g, out, write_full);  \
        break;
    bool done = true;
    switch (M) {
        LC3D_SOLVE_STATIC(128)
        LC3D_SOLVE_STATIC(256)
        LC3D_SOLVE_STATIC(320)
        LC3D_SOLVE_STATIC(384)
        LC3D_SOLVE_STATIC(512)
        LC3D_SOLVE_STATIC(768)
        LC3D_SOLVE_STATIC(1024)
    default: done = false;
    }
#undef LC3D_SOLVE_STATIC
    if (done) {
    } else if (M <= 64) {
        k_rows_fwd<TG, TPRED, 64><<<rows, 256, lds, s>>>(g, in, rows.x);
        k_cols<TG, 64><<<rows, 256, lds, s>>>(g);
        k_rows_inv<TG, TFULL, THALF, 64><<<rows, 256, lds, s>>>(g, out, write_full);
    } else if (M <= 255) {
        k_rows_fwd<TG, TPRED, 255><<<rows, 256, lds, s>>>(g, in, rows.x);
        k_cols<TG, 255><<<rows, 256, lds, s>>>(g);
        k_rows_inv<TG, TFULL, THALF, 255><<<rows, 256, lds, s>>>(g, out, write_full);
    } else if (M <= 511) {
        k_rows_fwd<TG, TPRED, 511><<<rows, 256, lds, s>>>(g, in, rows.x);
        k_cols<TG, 511><<<rows, 256, lds, s>>>(g);
        k_rows_inv<TG, TFULL, THALF, 511><<<rows, 256, lds, s>>>(g, out, write_full);
    } else {
        k_rows_fwd<TG, TPRED, 1023><<<rows, 256, lds, s>>>(g, in, rows.x);
        k_cols<TG, 1023><<<rows, 256, lds, s>>>(g);
        k_rows_inv<TG, TFULL, THALF, 1023><<<rows, 256, lds, s>>>(g, out, write_full);
    }
    const int fb = (int)std::min<int64_t>(1024, ((int64_t)N * N + 255) / 256);
    k_bz_finalize<TFULL, THALF><<<fb, 256, 0, s>>>(N, g.rowsum, out.prev[F_BZ], out.full[F_BZ],
                                                   out.half[F_BZ]);
    return hipGetLastError();
}

template hipError_t launch_compute_fields<double, double, double, double>(
    const SolveGeom &, const SolveIn<double, double> &, const SolveOut<double, double> &, bool,
    hipStream_t);
template hipError_t launch_compute_fields<float, double, double, float>(
    const SolveGeom &, const SolveIn<float, double> &, const SolveOut<double, float> &, bool,
    hipStream_t);
template hipError_t launch_compute_fields<float, float, double, float>(
    const SolveGeom &, const SolveIn<float, float> &, const SolveOut<double, float> &, bool,
    hipStream_t);

}  // namespace lc3d
