// Field solve of one xi step (plasma3d/fields.py:112-322) with every DST-I /
// DCT-I computed by a complex FFT of length L = 2(N-1) in LDS, gfx950.
//
// scipy.fftpack's DST-I and DCT-I of x_0..x_M (M = N - 1; the DST reads
// x_1..x_{M-1}), the transforms of fields.py:15-109, are the length-L DFT of
// x's odd / even period-L extension:
//     DCT-I(x)_k =  Re DFT(even extension)_k   (that DFT is real),
//     DST-I(x)_k = -Im DFT(odd extension)_k    (that DFT is imaginary).
// Two sequences a, b of one kind therefore share one complex FFT of
// z = ext(a) + i ext(b), Z_k = A_k + i B_k:
//     DCT: Y^a_k =  Re Z_k,  Y^b_k = Im Z_k;
//     DST: Y^a_k = -Im Z_k,  Y^b_k = Re Z_k.
// Each transform is its own inverse up to the factor L, so the solve of
// field f with right-hand side R is Ty Tx (Lambda .* (Tx Ty R)), Lambda the
// solver matrix of initialization.py:15-67 divided by L^2 (solver_factor).
//
// Kernels (all seven fields in one launch; rows / columns paired (2p, 2p + 1),
// the last pair of the odd N holding one; a 1-D grid whose blocks are dealt
// to the XCDs so that blocks reading the same rows or cache lines share an
// L2, see rf_block):
//   k_rf_rows_fwd  right-hand sides of two rows (evaluated here from the
//                  currents), transform along y, spectral rows into work;
//   k_rf_cols      two spectral columns: transform along x, times Lambda,
//                  transform back along x, in place in work;
//   k_rf_rows_inv  two rows back along y; outputs (fields.py:299-314) and the
//                  Bz row sums;
//   k_bz_finalize  removes the Bz mean (fields.py:264).
// The FFT is a Stockham autosort FFT on the tpb transforms of a block at
// once, each stage from one LDS buffer into the other (one barrier per
// stage).  The radix sequence is a host plan packed into one integer;
// twiddles are a host-built float64 table of exp(-2 pi i t / L) in the
// solve's type.
#include "solve_common.h"

// Index checks of every LDS and work-array access, compiled only into the
// host build of these kernels (tests/cpu_emu defines RF_CHECK).
#ifndef RF_CHECK
#define RF_CHECK(cond) ((void)0)
#endif

namespace lc3d {

constexpr int RF_T = 256;   // threads per block

template <typename T> struct Cx {
    using C = cplx_t<T>;
    static __device__ __forceinline__ C make(T re, T im) { C r; r.x = re; r.y = im; return r; }
    static __device__ __forceinline__ C add(C a, C b) { return make(a.x + b.x, a.y + b.y); }
    static __device__ __forceinline__ C sub(C a, C b) { return make(a.x - b.x, a.y - b.y); }
    static __device__ __forceinline__ C mul(C a, C b) {
        return make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
    static __device__ __forceinline__ C mul_c(C a, T re, T im) {   // a * (re + i im)
        return make(a.x * re - a.y * im, a.x * im + a.y * re);
    }
    static __device__ __forceinline__ C negi(C a) { return make(a.y, -a.x); }   // a * (-i)
    static __device__ __forceinline__ C neg(C a) { return make(-a.x, -a.y); }
};

// x / d and x mod d for 0 <= x < 2^22, d >= 1, by one float multiply: the
// true quotient (x + 1/2) / d is at least 1/(2d) from an integer, and the
// float product's relative error (< 2^-23) stays below that for such x.
struct RfDiv {
    int d;
    float inv;
    __device__ __forceinline__ explicit RfDiv(int d_) : d(d_), inv(1.0f / (float)d_) {}
    __device__ __forceinline__ int div(int x) const
    {
        const int q = (int)(((float)x + 0.5f) * inv);
        RF_CHECK(x >= 0 && x < (1 << 22) && q == x / d);
        return q;
    }
};

// Forward 4-point DFT in place.
template <typename T>
__device__ __forceinline__ void rf_dft4(cplx_t<T> &a0, cplx_t<T> &a1, cplx_t<T> &a2, cplx_t<T> &a3)
{
    using X = Cx<T>;
    const cplx_t<T> s02 = X::add(a0, a2), d02 = X::sub(a0, a2);
    const cplx_t<T> s13 = X::add(a1, a3), d13 = X::negi(X::sub(a1, a3));
    a0 = X::add(s02, s13);
    a1 = X::add(d02, d13);
    a2 = X::sub(s02, s13);
    a3 = X::sub(d02, d13);
}

// R-point forward DFT y_q = sum_p v_p exp(-2 pi i p q / R) in place.  The
// general form (odd R) reads exp(-2 pi i m / R) from the twiddle table.
template <typename T, int R> struct RfDft {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[R], const cplx_t<T> *__restrict__ tw,
                                               int L)
    {
        using X = Cx<T>;
        const int step = L / R;
        cplx_t<T> y[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cplx_t<T> s = v[0];
#pragma unroll
            for (int p = 1; p < R; ++p)
                s = (q == 0) ? X::add(s, v[p]) : X::add(s, X::mul(v[p], tw[((p * q) % R) * step]));
            y[q] = s;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = y[q];
    }
};

template <typename T> struct RfDft<T, 2> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[2], const cplx_t<T> *, int)
    {
        const cplx_t<T> a = v[0], b = v[1];
        v[0] = Cx<T>::add(a, b);
        v[1] = Cx<T>::sub(a, b);
    }
};

template <typename T> struct RfDft<T, 3> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[3], const cplx_t<T> *, int)
    {
        using X = Cx<T>;
        const T h3 = T(0.86602540378443864676);   // sqrt(3) / 2
        const cplx_t<T> t = X::add(v[1], v[2]);
        const cplx_t<T> d = X::negi(X::sub(v[1], v[2]));   // -i (v1 - v2)
        const cplx_t<T> m = X::make(v[0].x - T(0.5) * t.x, v[0].y - T(0.5) * t.y);
        v[0] = X::add(v[0], t);
        v[1] = X::make(m.x + h3 * d.x, m.y + h3 * d.y);
        v[2] = X::make(m.x - h3 * d.x, m.y - h3 * d.y);
    }
};

template <typename T> struct RfDft<T, 4> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[4], const cplx_t<T> *, int)
    {
        rf_dft4<T>(v[0], v[1], v[2], v[3]);
    }
};

// 8 = 2 x 4: DFTs of the even and odd samples, combined with exp(-i pi k / 4).
template <typename T> struct RfDft<T, 8> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[8], const cplx_t<T> *, int)
    {
        using X = Cx<T>;
        const T r2 = T(0.70710678118654752440);   // sqrt(1/2)
        cplx_t<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        cplx_t<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        rf_dft4<T>(e0, e1, e2, e3);
        rf_dft4<T>(o0, o1, o2, o3);
        o1 = X::make(r2 * (o1.x + o1.y), r2 * (o1.y - o1.x));    // * exp(-i pi / 4)
        o2 = X::negi(o2);                                         // * exp(-i pi / 2)
        o3 = X::make(r2 * (o3.y - o3.x), -r2 * (o3.x + o3.y));   // * exp(-3 i pi / 4)
        v[0] = X::add(e0, o0);
        v[4] = X::sub(e0, o0);
        v[1] = X::add(e1, o1);
        v[5] = X::sub(e1, o1);
        v[2] = X::add(e2, o2);
        v[6] = X::sub(e2, o2);
        v[3] = X::add(e3, o3);
        v[7] = X::sub(e3, o3);
    }
};

// 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.  DFT4 over n1 for each n2, times
// w^(n2 k1) (w = exp(-2 pi i / 16)), DFT4 over n2 for each k1.
template <typename T> struct RfDft<T, 16> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[16], const cplx_t<T> *, int)
    {
        using X = Cx<T>;
        const T c8 = T(0.92387953251128675613);   // cos(pi / 8)
        const T s8 = T(0.38268343236508977173);   // sin(pi / 8)
        const T r2 = T(0.70710678118654752440);
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) rf_dft4<T>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
        // v[4 k1 + n2] now holds the k1-th output of group n2
        v[4 + 1] = X::mul_c(v[4 + 1], c8, -s8);     // w^1
        v[4 + 2] = X::mul_c(v[4 + 2], r2, -r2);     // w^2
        v[4 + 3] = X::mul_c(v[4 + 3], s8, -c8);     // w^3
        v[8 + 1] = X::mul_c(v[8 + 1], r2, -r2);     // w^2
        v[8 + 2] = X::negi(v[8 + 2]);               // w^4
        v[8 + 3] = X::mul_c(v[8 + 3], -r2, -r2);    // w^6
        v[12 + 1] = X::mul_c(v[12 + 1], s8, -c8);   // w^3
        v[12 + 2] = X::mul_c(v[12 + 2], -r2, -r2);  // w^6
        v[12 + 3] = X::mul_c(v[12 + 3], -c8, s8);   // w^9
        cplx_t<T> y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cplx_t<T> b0 = v[4 * k1], b1 = v[4 * k1 + 1], b2 = v[4 * k1 + 2], b3 = v[4 * k1 + 3];
            rf_dft4<T>(b0, b1, b2, b3);
            y[k1] = b0;
            y[k1 + 4] = b1;
            y[k1 + 8] = b2;
            y[k1 + 12] = b3;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = y[k];
    }
};

// 32 = 2 x 16: DFTs of the even and odd samples, combined with
// exp(-2 pi i k / 32).
template <typename T> struct RfDft<T, 32> {
    static __device__ __forceinline__ void run(cplx_t<T> (&v)[32], const cplx_t<T> *tw, int L)
    {
        using X = Cx<T>;
        // cos and sin of pi k / 16, k = 0 .. 8
        const T c[9] = {T(1.0),
                        T(0.98078528040323044913),
                        T(0.92387953251128675613),
                        T(0.83146961230254523708),
                        T(0.70710678118654752440),
                        T(0.55557023301960222474),
                        T(0.38268343236508977173),
                        T(0.19509032201612826785),
                        T(0.0)};
        cplx_t<T> e[16], o[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        RfDft<T, 16>::run(e, tw, L);
        RfDft<T, 16>::run(o, tw, L);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            // exp(-i pi k / 16) = cos - i sin; sin(pi k / 16) = cos(pi (8 - k) / 16)
            const T cr = k <= 8 ? c[k] : -c[16 - k];
            const T si = k <= 8 ? c[8 - k] : c[k - 8];
            const cplx_t<T> t = X::mul_c(o[k], cr, -si);
            v[k] = X::add(e[k], t);
            v[k + 16] = X::sub(e[k], t);
        }
    }
};

// ---- one transform per wave, in registers ----------------------------------------
// Length L = 64 P (P = 8, 16, 32 for L = 512, 1024, 2048), R3 = 64 / P.  With
// n = 64 r + t and k = q + P k' the DFT is a radix-P DFT over r for each lane
// t, twiddled by w_L^(t q), then for each q a 64-point DFT over t, itself a
// radix-P DFT over s (t = u + R3 s) twiddled by w_64^(u m), and a radix-R3
// DFT over u (k' = m + P w):
//   stage 1  lane t        reads x[64 r + t]            writes B[q S1 + t]
//   stage 2  lane q R3 + u reads B[q S1 + u + R3 s]     writes C[u S2 + m P + q]
//   stage 3  lane q + P mm reads C[u S2 + m P + q]      writes X[q + P m + P^2 w]
//            for m = mm + R3 j, j < P / R3
// in the transform's own LDS region of L + 64 values, S1 = 64 + R3 and
// S2 = P^2 + P, strides under which every access of a wave is spread over
// all banks.  The block's four waves take its transforms in turn; the
// result is in natural order in place.
constexpr int RF_WAVE_TPB = RF_T / 64;   // transforms per block, one per wave
template <int L> struct RfWave {
    static constexpr int P = L / 64, R3 = 64 / P, RS = L + 64;
    static constexpr int S1 = 64 + R3, S2 = P * P + P;
    static_assert(P * R3 == 64 && P >= R3 && P * S1 <= RS && R3 * S2 <= RS, "wave FFT shape");
};

template <typename T, int L>
__device__ __forceinline__ void rf_fft_wave(cplx_t<T> *lds, const cplx_t<T> *__restrict__ tw,
                                            int ntr)
{
    using C = cplx_t<T>;
    using X = Cx<T>;
    using W = RfWave<L>;
    constexpr int P = W::P, R3 = W::R3;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rounds = (ntr + RF_WAVE_TPB - 1) / RF_WAVE_TPB;
    for (int round = 0; round < rounds; ++round) {
        const int tr = round * RF_WAVE_TPB + wave;
        const bool act = tr < ntr;
        C *x = lds + (act ? tr : 0) * W::RS;
        C v[P];
        // stage 1
        if (act) {
#pragma unroll
            for (int r = 0; r < P; ++r) {
                RF_CHECK(64 * r + lane < W::RS);
                v[r] = x[64 * r + lane];
            }
            RfDft<T, P>::run(v, tw, L);
#pragma unroll
            for (int q = 1; q < P; ++q) v[q] = X::mul(v[q], tw[lane * q]);
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < P; ++q) x[q * W::S1 + lane] = v[q];
        }
        __syncthreads();
        // stage 2
        const int q2 = lane / R3, u = lane % R3;
        if (act) {
#pragma unroll
            for (int s = 0; s < P; ++s) {
                RF_CHECK(q2 * W::S1 + u + R3 * s < W::RS);
                v[s] = x[q2 * W::S1 + u + R3 * s];
            }
            RfDft<T, P>::run(v, tw, L);
#pragma unroll
            for (int m = 1; m < P; ++m) v[m] = X::mul(v[m], tw[P * u * m]);
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                RF_CHECK(u * W::S2 + m * P + q2 < W::RS);
                x[u * W::S2 + m * P + q2] = v[m];
            }
        }
        __syncthreads();
        // stage 3
        const int q3 = lane % P, mm = lane / P;
        if (act) {
#pragma unroll
            for (int j = 0; j < P / R3; ++j) {
                C w[R3];
#pragma unroll
                for (int uu = 0; uu < R3; ++uu) w[uu] = x[uu * W::S2 + (mm + R3 * j) * P + q3];
                RfDft<T, R3>::run(w, tw, L);
#pragma unroll
                for (int ww = 0; ww < R3; ++ww) v[j * R3 + ww] = w[ww];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int j = 0; j < P / R3; ++j)
#pragma unroll
                for (int ww = 0; ww < R3; ++ww) {
                    const int k = q3 + P * (mm + R3 * j) + P * P * ww;
                    RF_CHECK(k < L);
                    x[k] = v[j * R3 + ww];
                }
        }
        __syncthreads();
    }
}

// One Stockham stage of radix R over the ntr transforms of length L in src
// (transform t at src[t L ..]) into dst; Ns is the product of the earlier
// radices.  Butterfly b of a transform reads x[b + r L/R], twiddles input r
// by exp(-2 pi i r k / (Ns R)), k = b mod Ns, and writes
// y[(b - k) R + k + r Ns].  Every index stays below L: Ns R divides L.
template <typename T, int R>
__device__ __forceinline__ void rf_stage(const cplx_t<T> *src, cplx_t<T> *dst,
                                         const cplx_t<T> *__restrict__ tw, int L, int ntr, int Ns)
{
    using C = cplx_t<T>;
    using X = Cx<T>;
    const int LR = L / R, total = ntr * LR, tws = L / (Ns * R);
    const RfDiv dLR(LR), dNs(Ns);
    for (int g = threadIdx.x; g < total; g += RF_T) {
        const int tr = dLR.div(g), b = g - tr * LR, k = b - dNs.div(b) * Ns;
        const C *in = src + tr * L + b;
        RF_CHECK(tr < ntr && b + (R - 1) * LR < L);
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[r * LR];
        if (k > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = X::mul(v[r], tw[r * k * tws]);
        }
        RfDft<T, R>::run(v, tw, L);
        C *out = dst + tr * L + (b - k) * R + k;
        RF_CHECK((b - k) * R + k + (R - 1) * Ns < L && (R - 1) * k * tws < L);
#pragma unroll
        for (int r = 0; r < R; ++r) out[r * Ns] = v[r];
    }
    __syncthreads();
}

// Forward FFT of the ntr length-L transforms in a, radices from `stages`,
// alternating between the LDS buffers a and b; returns the one holding the
// result.  Every thread of the block runs every stage (each ends with a
// barrier).
template <typename T>
__device__ __forceinline__ cplx_t<T> *rf_fft(cplx_t<T> *a, cplx_t<T> *b,
                                             const cplx_t<T> *__restrict__ tw, uint64_t stages,
                                             int L, int ntr)
{
    int Ns = 1;
    for (int s = 0; s < RFFT_MAX_STAGES; ++s) {
        const int R = (int)((stages >> (5 * s)) & 31);
        if (R == 0) break;
        switch (R) {
        case 2: rf_stage<T, 2>(a, b, tw, L, ntr, Ns); break;
        case 3: rf_stage<T, 3>(a, b, tw, L, ntr, Ns); break;
        case 4: rf_stage<T, 4>(a, b, tw, L, ntr, Ns); break;
        case 5: rf_stage<T, 5>(a, b, tw, L, ntr, Ns); break;
        case 7: rf_stage<T, 7>(a, b, tw, L, ntr, Ns); break;
        case 8: rf_stage<T, 8>(a, b, tw, L, ntr, Ns); break;
        case 11: rf_stage<T, 11>(a, b, tw, L, ntr, Ns); break;
        case 13: rf_stage<T, 13>(a, b, tw, L, ntr, Ns); break;
        case 16: rf_stage<T, 16>(a, b, tw, L, ntr, Ns); break;
        default: break;   // the host accepts only the radices above
        }
        cplx_t<T> *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// Transforms (row or column pairs) of block bx: min(tpb, pairs - bx tpb).
__device__ __forceinline__ int rf_block_transforms(int tpb, int pairs, int bx)
{
    const int left = pairs - bx * tpb;
    return left < tpb ? left : tpb;
}

// Field and pair group of this block.  The hardware deals consecutive block
// ids round-robin over the 8 XCDs; xcd_swizzle turns that into contiguous
// runs of logical ids per XCD, in launch order.  With fields_fastest the
// seven fields of a row group are consecutive (the row passes: they read the
// same current rows), otherwise the groups of a field are (the column pass:
// neighbouring column groups share cache lines).
__device__ __forceinline__ void rf_block(int groups, bool fields_fastest, int &f, int &group)
{
    const int total = 7 * groups;
    const int l = xcd_swizzle((int)blockIdx.x, total);
    if (fields_fastest) {
        group = l / 7;
        f = l - 7 * group;
    } else {
        f = l / groups;
        group = l - f * groups;
    }
    RF_CHECK(f >= 0 && f < 7 && group >= 0 && group < groups);
}



// ---- rows, forward: right-hand sides and the transform along y ---------------
// CAP: complex values of the block's LDS, two buffers of CAP / 2.
// LDS of a block: CAP complex values for the Stockham FFT, four regions of
// WL + 64 for the FFT by waves of length WL.
template <int CAP, int WL> struct RfLds {
    static constexpr int size = WL > 0 ? RF_WAVE_TPB * (WL + 64) : CAP;
};

// Runs the forward FFT of the block's ntr transforms, stored from `a` (region
// stride L, or WL + 64 for the wave FFT); returns where the result is (the
// Stockham FFT alternates between a and b, the wave FFT works in place).
template <typename TG, int WL>
__device__ __forceinline__ cplx_t<TG> *rf_transform(cplx_t<TG> *a, cplx_t<TG> *b,
                                                    const cplx_t<TG> *__restrict__ tw,
                                                    uint64_t stages, int L, int ntr)
{
    if constexpr (WL > 0) {
        rf_fft_wave<TG, WL>(a, tw, ntr);
        return a;
    } else {
        return rf_fft<TG>(a, b, tw, stages, L, ntr);
    }
}

template <typename TG, typename TPRED, int CAP, int WL>
__global__ __launch_bounds__(RF_T) void k_rf_rows_fwd(RfftGeom g, RhsEval<TG, TPRED> rhs)
{
    using C = cplx_t<TG>;
    using X = Cx<TG>;
    constexpr int LDS = RfLds<CAP, WL>::size;
    __shared__ __align__(16) C lds[LDS];
    const int N = g.N, M = N - 1, L = 2 * M, per = M + 1, groups = (N + 1) / 2;
    int f, grp;
    const int tpb = WL > 0 ? RF_WAVE_TPB : g.tpb;
    rf_block((groups + tpb - 1) / tpb, true, f, grp);
    const int p0 = grp * tpb;
    const int ntr = rf_block_transforms(tpb, groups, grp);
    const bool ydst = dst_along_y(f);
    const C *tw = static_cast<const C *>(g.tw);
    TG *work = static_cast<TG *>(g.work) + (size_t)f * N * N;
    const RfDiv dper(per);
    const int RS = WL > 0 ? WL + 64 : L;   // region stride of a transform
    RF_CHECK(ntr >= 1 && (WL > 0 ? (WL == L && ntr <= RF_WAVE_TPB) : 2 * ntr * L <= CAP));
    for (int it = threadIdx.x; it < ntr * per; it += RF_T) {
        const int tr = dper.div(it), j = it - tr * per;
        const int ia = 2 * (p0 + tr), ib = ia + 1;
        RF_CHECK(ia < N && j <= M);
        const TG a = in_rhs_domain(f, ia, j, N) ? rhs(f, ia, j) : TG(0);
        const TG b = (ib < N && in_rhs_domain(f, ib, j, N)) ? rhs(f, ib, j) : TG(0);
        C *t = lds + tr * RS;
        t[j] = X::make(a, b);
        if (j > 0 && j < M) t[L - j] = ydst ? X::make(-a, -b) : X::make(a, b);
    }
    __syncthreads();
    const C *z = rf_transform<TG, WL>(lds, lds + LDS / 2, tw, g.stages, L, ntr);
    for (int it = threadIdx.x; it < ntr * per; it += RF_T) {
        const int tr = dper.div(it), k = it - tr * per;
        const int ia = 2 * (p0 + tr), ib = ia + 1;
        const C v = z[tr * RS + k];
        const bool zero = ydst && (k == 0 || k == M);
        const TG ya = zero ? TG(0) : (ydst ? -v.y : v.x);
        const TG yb = zero ? TG(0) : (ydst ? v.x : v.y);
        work[(size_t)ia * N + k] = ya;
        if (ib < N) work[(size_t)ib * N + k] = yb;
    }
}

// ---- columns: along x, times the solver matrix, back along x -----------------
template <typename TG, int CAP, int WL>
__global__ __launch_bounds__(RF_T) void k_rf_cols(RfftGeom g)
{
    using C = cplx_t<TG>;
    using X = Cx<TG>;
    constexpr int LDS = RfLds<CAP, WL>::size;
    __shared__ __align__(16) C lds[LDS];
    const int N = g.N, M = N - 1, L = 2 * M, per = M + 1, groups = (N + 1) / 2;
    int f, grp;
    const int tpb = WL > 0 ? RF_WAVE_TPB : g.tpb;
    rf_block((groups + tpb - 1) / tpb, false, f, grp);
    const int p0 = grp * tpb;
    const int ntr = rf_block_transforms(tpb, groups, grp);
    const int j0 = 2 * p0, nc = 2 * ntr;   // columns j0 .. j0 + nc - 1 (the last may be N)
    const RfDiv dper(per), dnc(nc);
    const bool xdst = dst_along_x(f);
    const C *tw = static_cast<const C *>(g.tw);
    TG *work = static_cast<TG *>(g.work) + (size_t)f * N * N;
    C *a = lds, *b = lds + LDS / 2;
    TG *ar = reinterpret_cast<TG *>(a);
    const int RS = WL > 0 ? WL + 64 : L;   // region stride of a transform
    RF_CHECK(ntr >= 1 && (WL > 0 ? (WL == L && ntr <= RF_WAVE_TPB) : 2 * ntr * L <= CAP) &&
             j0 + nc <= N + 1);
    // Consecutive threads take consecutive columns of one row.
    for (int it = threadIdx.x; it < per * nc; it += RF_T) {
        const int i = dnc.div(it), c = it - i * nc;
        const int j = j0 + c, tr = c >> 1, part = c & 1;
        TG x = j < N ? work[(size_t)i * N + j] : TG(0);
        if (xdst && (i == 0 || i == M)) x = TG(0);
        ar[2 * (tr * RS + i) + part] = x;
        if (i > 0 && i < M) ar[2 * (tr * RS + L - i) + part] = xdst ? -x : x;
    }
    __syncthreads();
    C *z = rf_transform<TG, WL>(a, b, tw, g.stages, L, ntr);
    // The wave FFT works in place: the scaled spectra overwrite it (a thread
    // writes the k it read, and L - k > M, which no thread reads).
    C *e = WL > 0 ? z : (z == a ? b : a);
    // Spectra times the solver matrix (zero at DST edges), written as the
    // extension the inverse transform along x starts from.
    const ScaleGeom sg{g.lam, g.c, (double)L * (double)L};
    for (int it = threadIdx.x; it < ntr * per; it += RF_T) {
        const int tr = dper.div(it), k = it - tr * per;
        const int ja = j0 + 2 * tr, jb = ja + 1;
        const C v = z[tr * RS + k];
        const TG ya = xdst ? -v.y : v.x;
        const TG yb = xdst ? v.x : v.y;
        const double fa = solver_factor(f, k, ja, N, sg);
        const double fb = jb < N ? solver_factor(f, k, jb, N, sg) : 0.0;
        const C s = X::make((TG)((double)ya * fa), (TG)((double)yb * fb));
        C *t = e + tr * RS;
        t[k] = s;
        if (k > 0 && k < M) t[L - k] = xdst ? X::neg(s) : s;
    }
    __syncthreads();
    z = rf_transform<TG, WL>(e, z, tw, g.stages, L, ntr);
    for (int it = threadIdx.x; it < per * nc; it += RF_T) {
        const int i = dnc.div(it), c = it - i * nc;
        const int j = j0 + c, tr = c >> 1, part = c & 1;
        if (j < N) {
            const C v = z[tr * RS + i];
            TG x;
            if (xdst) x = (i == 0 || i == M) ? TG(0) : (part ? v.x : -v.y);
            else x = part ? v.y : v.x;
            work[(size_t)i * N + j] = x;
        }
    }
}

// ---- rows, inverse: back along y, outputs and the Bz row sums ----------------
template <typename TG, typename TFULL, typename THALF, int CAP, int WL>
__global__ __launch_bounds__(RF_T) void k_rf_rows_inv(RfftGeom g, SolveOut<TFULL, THALF> out,
                                                      bool write_full)
{
    using C = cplx_t<TG>;
    using X = Cx<TG>;
    constexpr int LDS = RfLds<CAP, WL>::size;
    __shared__ __align__(16) C lds[LDS];
    const int N = g.N, M = N - 1, L = 2 * M, per = M + 1, groups = (N + 1) / 2;
    int f, grp;
    const int tpb = WL > 0 ? RF_WAVE_TPB : g.tpb;
    rf_block((groups + tpb - 1) / tpb, true, f, grp);
    const int p0 = grp * tpb;
    const int ntr = rf_block_transforms(tpb, groups, grp);
    const bool ydst = dst_along_y(f);
    const C *tw = static_cast<const C *>(g.tw);
    const TG *work = static_cast<const TG *>(g.work) + (size_t)f * N * N;
    const RfDiv dper(per);
    const int RS = WL > 0 ? WL + 64 : L;   // region stride of a transform
    RF_CHECK(ntr >= 1 && (WL > 0 ? (WL == L && ntr <= RF_WAVE_TPB) : 2 * ntr * L <= CAP));
    for (int it = threadIdx.x; it < ntr * per; it += RF_T) {
        const int tr = dper.div(it), k = it - tr * per;
        const int ia = 2 * (p0 + tr), ib = ia + 1;
        RF_CHECK(ia < N && k <= M);
        TG a = work[(size_t)ia * N + k];
        TG b = ib < N ? work[(size_t)ib * N + k] : TG(0);
        if (ydst && (k == 0 || k == M)) a = b = TG(0);
        C *t = lds + tr * RS;
        t[k] = X::make(a, b);
        if (k > 0 && k < M) t[L - k] = ydst ? X::make(-a, -b) : X::make(a, b);
    }
    __syncthreads();
    const C *z = rf_transform<TG, WL>(lds, lds + LDS / 2, tw, g.stages, L, ntr);
    for (int it = threadIdx.x; it < ntr * per; it += RF_T) {
        const int tr = dper.div(it), j = it - tr * per;
        const int ia = 2 * (p0 + tr), ib = ia + 1;
        const C v = z[tr * RS + j];
        const bool zero = ydst && (j == 0 || j == M);
        const TG xa = zero ? TG(0) : (ydst ? -v.y : v.x);
        const TG xb = zero ? TG(0) : (ydst ? v.x : v.y);
        emit_output(out, f, (size_t)ia * N + j, xa, write_full);
        if (ib < N) emit_output(out, f, (size_t)ib * N + j, xb, write_full);
    }
    if (f == F_BZ) {
        // Row sums of the raw Bz in a fixed order: one wave per row, lanes
        // over j, a butterfly reduction (z is not written after the FFT).
        const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
        for (int r = wave; r < 2 * ntr; r += RF_T / 64) {
            const int tr = r >> 1, part = r & 1, row = 2 * (p0 + tr) + part;
            double s = 0;
            for (int j = lane; j <= M; j += 64) {
                const C v = z[tr * RS + j];
                const bool zero = ydst && (j == 0 || j == M);
                const TG x = zero ? TG(0) : (part ? (ydst ? v.x : v.y) : (ydst ? -v.y : v.x));
                s += (double)x;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0 && row < N) g.rowsum[row] = s;
        }
    }
}

template <typename TG> struct RfCaps;
// 16, 32 and 64 KB of LDS per block.
template <> struct RfCaps<float> { static constexpr int c0 = 2048, c1 = 4096, c2 = 8192; };
template <> struct RfCaps<double> { static constexpr int c0 = 1024, c1 = 2048, c2 = 4096; };

bool rfft_cap_supported(bool fp32, int cap)
{
    return fp32 ? (cap == RfCaps<float>::c0 || cap == RfCaps<float>::c1 || cap == RfCaps<float>::c2)
                : (cap == RfCaps<double>::c0 || cap == RfCaps<double>::c1 ||
                   cap == RfCaps<double>::c2);
}

// Launches one of the three kernels, by waves (wave-FFT length WL) or by the
// Stockham FFT in a block of CAP complex values.
template <int K, typename TG, typename TPRED, typename TFULL, typename THALF, int CAP, int WL>
void rf_kernel(const RfftGeom &g, const RhsEval<TG, TPRED> &rhs,
               const SolveOut<TFULL, THALF> &out, bool write_full, hipStream_t s)
{
    const int tpb = WL > 0 ? RF_WAVE_TPB : g.tpb;
    const unsigned grid = 7 * (unsigned)(((g.N + 1) / 2 + tpb - 1) / tpb);
    if constexpr (K == 0) k_rf_rows_fwd<TG, TPRED, CAP, WL><<<grid, RF_T, 0, s>>>(g, rhs);
    else if constexpr (K == 1) k_rf_cols<TG, CAP, WL><<<grid, RF_T, 0, s>>>(g);
    else k_rf_rows_inv<TG, TFULL, THALF, CAP, WL><<<grid, RF_T, 0, s>>>(g, out, write_full);
}

template <int K, typename TG, typename TPRED, typename TFULL, typename THALF>
void rf_kernel_any(const RfftGeom &g, const RhsEval<TG, TPRED> &rhs,
                   const SolveOut<TFULL, THALF> &out, bool write_full, hipStream_t s)
{
    const int L = 2 * (g.N - 1);
    if (g.wave & (1 << K)) {
        if (L == 512) rf_kernel<K, TG, TPRED, TFULL, THALF, 0, 512>(g, rhs, out, write_full, s);
        else if (L == 1024) rf_kernel<K, TG, TPRED, TFULL, THALF, 0, 1024>(g, rhs, out, write_full, s);
        else rf_kernel<K, TG, TPRED, TFULL, THALF, 0, 2048>(g, rhs, out, write_full, s);
    } else if (g.cap == RfCaps<TG>::c0) {
        rf_kernel<K, TG, TPRED, TFULL, THALF, RfCaps<TG>::c0, 0>(g, rhs, out, write_full, s);
    } else if (g.cap == RfCaps<TG>::c1) {
        rf_kernel<K, TG, TPRED, TFULL, THALF, RfCaps<TG>::c1, 0>(g, rhs, out, write_full, s);
    } else {
        rf_kernel<K, TG, TPRED, TFULL, THALF, RfCaps<TG>::c2, 0>(g, rhs, out, write_full, s);
    }
}

bool rfft_wave_supported(int L) { return L == 512 || L == 1024 || L == 2048; }

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields_rfft(const RfftGeom &g, const SolveIn<TG, TPRED> &in,
                                      const SolveOut<TFULL, THALF> &out, bool write_full,
                                      hipStream_t s, KernelTimer *timer)
{
    const int N = g.N, L = 2 * (N - 1);
    // A kernel not by waves needs the Stockham plan and launch shape.
    if (g.wave < 0 || g.wave > 7 || (g.wave && !rfft_wave_supported(L)) ||
        (g.wave != 7 && (N < 9 || g.stages == 0 || g.tpb < 1 || 2 * (int64_t)g.tpb * L > g.cap ||
                         !rfft_cap_supported(sizeof(TG) == 4, g.cap))))
        return hipErrorInvalidValue;
    const RhsEval<TG, TPRED> rhs{in, N, TG(1.0 / (2.0 * g.h)), TG(1.0 / g.xi_step), TG(g.c)};
    hipError_t e;
    if (timer && (e = timer->begin(KC_SOLVE, 3, s)) != hipSuccess) return e;
    rf_kernel_any<0, TG, TPRED, TFULL, THALF>(g, rhs, out, write_full, s);
    rf_kernel_any<1, TG, TPRED, TFULL, THALF>(g, rhs, out, write_full, s);
    rf_kernel_any<2, TG, TPRED, TFULL, THALF>(g, rhs, out, write_full, s);
    if (timer && (e = timer->end(s)) != hipSuccess) return e;
    const int fb = (int)std::min<int64_t>(1024, ((int64_t)N * N + 255) / 256);
    k_bz_finalize<TFULL, THALF><<<fb, 256, 0, s>>>(N, g.rowsum, out.prev[F_BZ], out.full[F_BZ],
                                                   out.half[F_BZ]);
    return hipGetLastError();
}

template hipError_t launch_compute_fields_rfft<double, double, double, double>(
    const RfftGeom &, const SolveIn<double, double> &, const SolveOut<double, double> &, bool,
    hipStream_t, KernelTimer *);
template hipError_t launch_compute_fields_rfft<float, double, double, float>(
    const RfftGeom &, const SolveIn<float, double> &, const SolveOut<double, float> &, bool,
    hipStream_t, KernelTimer *);
template hipError_t launch_compute_fields_rfft<float, float, double, float>(
    const RfftGeom &, const SolveIn<float, float> &, const SolveOut<double, float> &, bool,
    hipStream_t, KernelTimer *);

}  // namespace lc3d
