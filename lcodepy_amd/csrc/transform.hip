// Field solve of one xi step (plasma3d/fields.py:112-322) with each DST-I /
// DCT-I written as a product with its N x N transform matrix, gfx950.
//
// For field f with right-hand side R (N x N, x index i = row, y index j =
// column), the solve is
//
//     X = Tx ( M .* (Tx R Ty^T) ) Ty^T
//
// where Tx, Ty are the DST-I matrix S or the DCT-I matrix C of the field's
// boundary kinds along x and y, and M holds 1/(l_kx + l_ky (+c)) / (2(N-1))^2
// (initialization.py:15-67; zero where a DST index is an edge, and at [0, 0]
// for the Neumann field Bz).  S and C are their own inverses up to the factor
// 2(N-1) that M carries, as the unnormalised transforms of fields.py:15-109.
// The edge rows/columns of S are zero, so DST directions see only the
// interior points and return exact zeros on the edges.
//
// Kernels, all on plain global buffers and static LDS:
//   k_rhs       R of all seven fields (one thread per grid point);
//   k_gemm_mfma C_f = A_f B_f for the seven fields of a launch (64 x 64
//               tiles through LDS, 16x16x4 matrix-core products), optionally
//               scaled by M;
//   k_assemble  outputs (full / half fields) and the Bz row sums;
//   k_bz_finalize  mean of Bz removed, half-step Bz.
#include "mfma.h"
#include "solve_common.h"

namespace lc3d {

template <typename T> struct GemmBatch {
    const T *A[7];
    const T *B[7];
    T *C[7];
};

struct ScaleGeom {
    const double *lam;
    double c;      // subtraction coefficient of the mixed fields
    double d2;     // (2 (N - 1))^2
};

template <typename TG, typename TPRED>
__global__ __launch_bounds__(256) void k_rhs(RhsEval<TG, TPRED> rhs, TG *__restrict__ work)
{
    const int N = rhs.N;
    const int f = blockIdx.y;
    const int64_t nn = (int64_t)N * N;
    const int64_t o = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (o >= nn) return;
    const int i = (int)(o / N), j = (int)(o % N);
    work[f * nn + o] = in_rhs_domain(f, i, j, N) ? rhs(f, i, j) : TG(0);
}

constexpr int GBM = 64, GBN = 64, GBK = 32;
constexpr int GLA = GBM * GBK / 256;   // A elements per thread per tile
constexpr int GLB = GBK * GBN / 256;   // B elements per thread per tile

// C_f = A_f B_f for the field f = blockIdx.z; all matrices N x N row-major.
// 256 threads = 4 waves; the block's 64 x 64 tile is split 2 x 2 between the
// waves, each wave computing 32 x 32 as 2 x 2 16x16x4 matrix-core products per
// step of 4 along k.  Tiles of A (64 x 32) and B (32 x 64) pass through
// static LDS; the next tiles' global loads are issued into registers before
// the current tiles are multiplied.  Loads past the matrix edge read as zero
// and stores past it are skipped.  With SCALE the result is multiplied by the
// solver matrix of the field at (row, column) = (kx, ky)
// (initialization.py:15-67).
template <typename T, bool SCALE>
__global__ __launch_bounds__(256) void k_gemm_mfma(GemmBatch<T> bt, int N, ScaleGeom sg)
{
    using MM = Mfma16<T>;
    using acc_t = typename MM::acc_t;
    __shared__ T As[GBM][GBK + 1];
    __shared__ T Bs[GBK][GBN + 1];
    const int f = blockIdx.z;
    const T *__restrict__ A = bt.A[f];
    const T *__restrict__ B = bt.B[f];
    T *__restrict__ C = bt.C[f];
    const int row0 = blockIdx.y * GBM, col0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;   // the wave's 32 x 32 in the tile
    const int li = lane % 16, lk = lane / 16;
    // Loader coordinates: A rows ar + 8 q, column ak; B rows bk + 4 q, column bc.
    const int ar = tid / GBK, ak = tid % GBK;
    const int bk = tid / GBN, bc = tid % GBN;
    acc_t acc[2][2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);

    T ra[GLA], rb[GLB];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < GLA; ++q) {
            const int gr = row0 + ar + 8 * q, gk = k0 + ak;
            ra[q] = (gr < N && gk < N) ? A[(size_t)gr * N + gk] : T(0);
        }
#pragma unroll
        for (int q = 0; q < GLB; ++q) {
            const int gk = k0 + bk + 4 * q, gc = col0 + bc;
            rb[q] = (gk < N && gc < N) ? B[(size_t)gk * N + gc] : T(0);
        }
    };
    load(0);
    for (int k0 = 0; k0 < N; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < GLA; ++q) As[ar + 8 * q][ak] = ra[q];
#pragma unroll
        for (int q = 0; q < GLB; ++q) Bs[bk + 4 * q][bc] = rb[q];
        __syncthreads();
        if (k0 + GBK < N) load(k0 + GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK; k4 += 4) {
            T a[2], b[2];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) a[ti] = As[wr + 16 * ti + li][k4 + lk];
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) b[tj] = Bs[k4 + lk][wc + 16 * tj + li];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = MM::mma(a[ti], b[tj], acc[ti][tj]);
        }
        __syncthreads();
    }

    const int M = N - 1;
    const bool dx = dst_along_x(f), dy = dst_along_y(f);
    const double cc = (f == F_EZ || f == F_BZ || f == F_PHI) ? 0.0 : sg.c;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = row0 + wr + 16 * ti + MM::row(lk, r);
                const int gc = col0 + wc + 16 * tj + li;
                if (gr < N && gc < N) {
                    T v = acc[ti][tj][r];
                    if (SCALE) {
                        // Spectral indices (kx, ky) = (gr, gc).
                        const bool ok = (!dx || (gr > 0 && gr < M)) && (!dy || (gc > 0 && gc < M)) &&
                                        !(f == F_BZ && gr == 0 && gc == 0);
                        const double mul = ok ? (1.0 / (sg.lam[gr] + sg.lam[gc] + cc)) / sg.d2 : 0.0;
                        v = (T)((double)v * mul);
                    }
                    C[(size_t)gr * N + gc] = v;
                }
            }
}

// One wave per (row, field): outputs of the row and, for Bz, its row sum.
template <typename TG, typename TFULL, typename THALF>
__global__ __launch_bounds__(256) void k_assemble(const TG *__restrict__ work, int N,
                                                  SolveOut<TFULL, THALF> out, bool write_full,
                                                  double *__restrict__ rowsum)
{
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int f = blockIdx.y;
    const int i = blockIdx.x * 4 + wave;
    if (i >= N) return;   // whole wave
    const size_t nn = (size_t)N * N;
    const TG *src = work + f * nn + (size_t)i * N;
    double s = 0;
    for (int j = lane; j < N; j += 64) {
        const TG r = src[j];
        emit_output(out, f, (size_t)i * N + j, r, write_full);
        s += (double)r;
    }
    if (f == F_BZ) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) rowsum[i] = s;
    }
}

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields_matrix(const MatrixGeom &g, const SolveIn<TG, TPRED> &in,
                                        const SolveOut<TFULL, THALF> &out, bool write_full,
                                        hipStream_t s, KernelTimer *timer)
{
    const int N = g.N;
    const size_t nn = (size_t)N * N;
    TG *w1 = static_cast<TG *>(g.work1);
    TG *w2 = static_cast<TG *>(g.work2);
    const TG *S = static_cast<const TG *>(g.tdst);
    const TG *C = static_cast<const TG *>(g.tdct);
    const TG *Ct = static_cast<const TG *>(g.tdct_t);
    const RhsEval<TG, TPRED> rhs{in, N, TG(1.0 / (2.0 * g.h)), TG(1.0 / g.xi_step), TG(g.c)};
    const ScaleGeom sg{g.lam, g.c, 4.0 * (double)(N - 1) * (double)(N - 1)};

    GemmBatch<TG> fy, sx, ix, iy;   // forward y, forward x (scaled), inverse x, inverse y
    for (int f = 0; f < 7; ++f) {
        const TG *Ty_t = dst_along_y(f) ? S : Ct;   // R Ty^T
        const TG *Tx = dst_along_x(f) ? S : C;      // Tx P
        fy.A[f] = w1 + f * nn; fy.B[f] = Ty_t; fy.C[f] = w2 + f * nn;
        sx.A[f] = Tx;          sx.B[f] = w2 + f * nn; sx.C[f] = w1 + f * nn;
        ix.A[f] = Tx;          ix.B[f] = w1 + f * nn; ix.C[f] = w2 + f * nn;
        iy.A[f] = w2 + f * nn; iy.B[f] = Ty_t; iy.C[f] = w1 + f * nn;
    }
    const dim3 tiles((N + GBN - 1) / GBN, (N + GBM - 1) / GBM, 7);
    k_rhs<TG, TPRED><<<dim3((unsigned)((nn + 255) / 256), 7), 256, 0, s>>>(rhs, w1);
    hipError_t e;
    if (timer && (e = timer->begin(KC_SOLVE_GEMM, 4, s)) != hipSuccess) return e;
    k_gemm_mfma<TG, false><<<tiles, 256, 0, s>>>(fy, N, sg);
    k_gemm_mfma<TG, true><<<tiles, 256, 0, s>>>(sx, N, sg);
    k_gemm_mfma<TG, false><<<tiles, 256, 0, s>>>(ix, N, sg);
    k_gemm_mfma<TG, false><<<tiles, 256, 0, s>>>(iy, N, sg);
    if (timer && (e = timer->end(s)) != hipSuccess) return e;
    k_assemble<TG, TFULL, THALF><<<dim3((N + 3) / 4, 7), 256, 0, s>>>(w1, N, out, write_full,
                                                                       g.rowsum);
    const int fb = (int)std::min<int64_t>(1024, ((int64_t)N * N + 255) / 256);
    k_bz_finalize<TFULL, THALF><<<fb, 256, 0, s>>>(N, g.rowsum, out.prev[F_BZ], out.full[F_BZ],
                                                   out.half[F_BZ]);
    return hipGetLastError();
}

template hipError_t launch_compute_fields_matrix<double, double, double, double>(
    const MatrixGeom &, const SolveIn<double, double> &, const SolveOut<double, double> &, bool,
    hipStream_t, KernelTimer *);
template hipError_t launch_compute_fields_matrix<float, double, double, float>(
    const MatrixGeom &, const SolveIn<float, double> &, const SolveOut<double, float> &, bool,
    hipStream_t, KernelTimer *);
template hipError_t launch_compute_fields_matrix<float, float, double, float>(
    const MatrixGeom &, const SolveIn<float, float> &, const SolveOut<double, float> &, bool,
    hipStream_t, KernelTimer *);

}  // namespace lc3d
