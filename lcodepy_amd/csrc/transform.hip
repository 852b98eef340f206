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
//   k_gemm_nn   C_f = A_f B_f for the seven fields of a launch (register-
//               tiled, 64 x 64 tiles through LDS), optionally scaled by M;
//   k_assemble  outputs (full / half fields) and the Bz row sums;
//   k_bz_finalize  mean of Bz removed, half-step Bz.
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

constexpr int GBM = 64, GBN = 64, GBK = 16;

// C_f = A_f B_f for the field f = blockIdx.z; all matrices N x N row-major.
// 256 threads: thread (tx, ty) = (tid % 16, tid / 16) owns rows ty + 16 r and
// columns tx + 16 c (r, c < 4) of the block's 64 x 64 tile.  Tiles of A
// (64 x 16, stored transposed) and B (16 x 64) pass through static LDS; loads
// past the matrix edge read as zero and stores past it are skipped.
template <typename T, bool SCALE>
__global__ __launch_bounds__(256) void k_gemm_nn(GemmBatch<T> bt, int N, ScaleGeom sg)
{
    __shared__ T As[GBK][GBM + 1];
    __shared__ T Bs[GBK][GBN];
    const int f = blockIdx.z;
    const T *__restrict__ A = bt.A[f];
    const T *__restrict__ B = bt.B[f];
    T *__restrict__ C = bt.C[f];
    const int row0 = blockIdx.y * GBM, col0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    T acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = T(0);

    for (int k0 = 0; k0 < N; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = ty + 16 * q;
            const int gr = row0 + r, gk = k0 + tx;
            As[tx][r] = (gr < N && gk < N) ? A[(size_t)gr * N + gk] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = tid / 64 + 4 * q, c = tid % 64;
            const int gk = k0 + kk, gc = col0 + c;
            Bs[kk][c] = (gk < N && gc < N) ? B[(size_t)gk * N + gc] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; ++kk) {
            T a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }

    const int M = N - 1;
    const bool dx = dst_along_x(f), dy = dst_along_y(f);
    const double cc = (f == F_EZ || f == F_BZ || f == F_PHI) ? 0.0 : sg.c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gr = row0 + ty + 16 * r;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gc = col0 + tx + 16 * c;
            if (gr < N && gc < N) {
                T v = acc[r][c];
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
                                        hipStream_t s)
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
    k_gemm_nn<TG, false><<<tiles, 256, 0, s>>>(fy, N, sg);
    k_gemm_nn<TG, true><<<tiles, 256, 0, s>>>(sx, N, sg);
    k_gemm_nn<TG, false><<<tiles, 256, 0, s>>>(ix, N, sg);
    k_gemm_nn<TG, false><<<tiles, 256, 0, s>>>(iy, N, sg);
    k_assemble<TG, TFULL, THALF><<<dim3((N + 3) / 4, 7), 256, 0, s>>>(w1, N, out, write_full,
                                                                       g.rowsum);
    const int fb = (int)std::min<int64_t>(1024, ((int64_t)N * N + 255) / 256);
    k_bz_finalize<TFULL, THALF><<<fb, 256, 0, s>>>(N, g.rowsum, out.prev[F_BZ], out.full[F_BZ],
                                                   out.half[F_BZ]);
    return hipGetLastError();
}

template hipError_t launch_compute_fields_matrix<double, double, double, double>(
    const MatrixGeom &, const SolveIn<double, double> &, const SolveOut<double, double> &, bool,
    hipStream_t);
template hipError_t launch_compute_fields_matrix<float, double, double, float>(
    const MatrixGeom &, const SolveIn<float, double> &, const SolveOut<double, float> &, bool,
    hipStream_t);
template hipError_t launch_compute_fields_matrix<float, float, double, float>(
    const MatrixGeom &, const SolveIn<float, float> &, const SolveOut<double, float> &, bool,
    hipStream_t);

}  // namespace lc3d
