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
//   k_rhs       R of all seven fields (one thread per grid point, all fields);
//   k_gemm_mfma C_f = A_f B_f for the seven fields of a launch (64 x 64
//               tiles through LDS, 16x16x4 matrix-core products; the last
//               row and column by the blocks of the last tile row and
//               column), optionally scaled by M;
//   k_assemble  outputs (full / half fields) and the Bz row sums;
//   k_bz_finalize  mean of Bz removed, half-step Bz.
#include "mfma.h"
#include <type_traits>
#include "solve_common.h"

namespace lc3d {

template <typename T> struct GemmBatch {
    const T *A[7];
    const T *B[7];
    T *C[7];
};


// One thread per grid point evaluates the seven fields' right-hand sides
// there (the field loop unrolled, so the currents they share are loaded
// once).
template <typename TG, typename TPRED>
__global__ __launch_bounds__(256) void k_rhs(RhsEval<TG, TPRED> rhs, TG *__restrict__ work)
{
    const int N = rhs.N;
    const int64_t nn = (int64_t)N * N;
    const int64_t o = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (o >= nn) return;
    const int i = (int)(o / N), j = (int)(o % N);
#pragma unroll
    for (int f = 0; f < 7; ++f) work[f * nn + o] = in_rhs_domain(f, i, j, N) ? rhs(f, i, j) : TG(0);
}

constexpr int GBM = 64, GBN = 64;   // output tile
constexpr int GTJ = GBN / 32;       // 16-column products per wave (waves 2 x 2)
// Depth of a k tile: 64 in float32, 32 in float64 (two blocks per CU fit
// the double-buffered LDS tiles either way).
template <typename T> struct GemmTile {
    static constexpr int BK = sizeof(T) == 4 ? 64 : 32;
    static constexpr int RS = BK + 4;            // LDS row stride of the A and B^T tiles
    static constexpr int NA = GBM * BK / 256;    // A elements per thread per tile
    static constexpr int NB = BK * GBN / 256;    // B elements per thread per tile
    static constexpr int RA = 256 / BK;          // A rows loaded per pass of the block
    static constexpr int KS = BK / 4;            // products per wave and 16 x 16 tile
};

// C_f = A_f B_f for the field f = blockIdx.z; all matrices N x N row-major.
// Over k < M = N - 1 the products run on the matrix cores in 64 x GBN output
// tiles covering [0, M) (rounded up to whole tiles); the k = M term is added
// in the epilogue.  Where M is a multiple of the tile size the last row
// (column) of index M lies just outside the tiles: the blocks of the last
// tile row also load row M of A into LDS with each k tile and accumulate row
// M of C for their columns on the vector units from the same B tile
// (likewise the last tile column for column M, and the corner block for
// C[M][M]).  Otherwise row (column) M lies inside the last tiles.
// 256 threads = 4 waves; a 64 x GBN tile is split 2 x 2 between the waves,
// each computing 32 x GBN/2 as 2 x GBN/32 16x16x4 matrix-core products.
// Tiles of A (64 x BK, row-major) and B (stored transposed, GBN x BK) pass
// through double-buffered static LDS, one barrier per tile (depth BK,
// GemmTile); a tile's global loads are issued two tiles before it is
// multiplied (two register sets).  Within a tile, step s (0..KS-1, KS =
// BK / 4) of a wave's products takes k = KS lk + s for lane group lk =
// lane / 16, the same for A and B, so each lane reads its KS A (and B)
// values of the tile as 16-byte LDS loads; the edge sums take the k slice
// KS (tid / 64) .. + KS - 1 per thread and add the four slices in a fixed
// order.  With SCALE the result is multiplied by the solver matrix of the
// field (solver_factor).
template <typename T, bool SCALE>
__global__ __launch_bounds__(256, 2) void k_gemm_mfma(GemmBatch<T> bt, int N, ScaleGeom sg)
{
    using MM = Mfma16<T>;
    using acc_t = typename MM::acc_t;
    constexpr int GBK = GemmTile<T>::BK, GRS = GemmTile<T>::RS, GLA = GemmTile<T>::NA,
                  GLB = GemmTile<T>::NB, GRA = GemmTile<T>::RA, KS = GemmTile<T>::KS;
    __shared__ __attribute__((aligned(16))) T As[2][GBM][GRS];
    __shared__ __attribute__((aligned(16))) T Bt[2][GBN][GRS];
    __shared__ T Xa[2][GBK];        // row M of A, this k tile
    __shared__ T Xb[2][GBK];        // column M of B, this k tile
    __shared__ T red[2][4][GBM];    // edge partial sums per k slice
    const int f = blockIdx.z;
    const T *__restrict__ A = bt.A[f];
    const T *__restrict__ B = bt.B[f];
    T *__restrict__ C = bt.C[f];
    const int M = N - 1;
    const int row0 = blockIdx.y * GBM, col0 = blockIdx.x * GBN;
    const bool xrow = M % GBM == 0 && row0 + GBM == M;   // this block also sums row M
    const bool xcol = M % GBN == 0 && col0 + GBN == M;   // ... and / or column M
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int wr = (wave / 2) * 32, wc = (wave % 2) * (GBN / 2);   // the wave's 32 x GBN/2
    const int li = lane % 16, lk = lane / 16;
    // Loader coordinates: A rows ar + GRA q, column ak; B rows bk + GRB q, column bc.
    const int ar = tid / GBK, ak = tid % GBK;
    constexpr int GRB = 256 / GBN;
    const int bk = tid / GBN, bc = tid % GBN;
    acc_t acc[2][GTJ];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < GTJ; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);
    T er = 0, ec = 0, ex = 0;   // row M / column M partial sums, corner

    // Two register sets: tile u is loaded into set u % 2, two tiles ahead of
    // the tile being multiplied.
    T ra[2][GLA], rb[2][GLB], xv[2] = {T(0), T(0)};
    auto load = [&](auto set, int k0) {
        constexpr int S = decltype(set)::value;
#pragma unroll
        for (int q = 0; q < GLA; ++q) {
            const int gr = row0 + ar + GRA * q, gk = k0 + ak;
            ra[S][q] = (gr < N && gk < M) ? A[(size_t)gr * N + gk] : T(0);
        }
#pragma unroll
        for (int q = 0; q < GLB; ++q) {
            const int gk = k0 + bk + GRB * q, gc = col0 + bc;
            rb[S][q] = (gk < M && gc < N) ? B[(size_t)gk * N + gc] : T(0);
        }
        if (xrow && tid < GBK) xv[S] = k0 + tid < M ? A[(size_t)M * N + k0 + tid] : T(0);
        if (xcol && tid >= 64 && tid < 64 + GBK) {
            const int gk = k0 + tid - 64;
            xv[S] = gk < M ? B[(size_t)gk * N + M] : T(0);
        }
    };
    auto store = [&](auto set) {   // set S into LDS buffer S
        constexpr int S = decltype(set)::value;
#pragma unroll
        for (int q = 0; q < GLA; ++q) As[S][ar + GRA * q][ak] = ra[S][q];
#pragma unroll
        for (int q = 0; q < GLB; ++q) Bt[S][bc][bk + GRB * q] = rb[S][q];
        if (xrow && tid < GBK) Xa[S][tid] = xv[S];
        if (xcol && tid >= 64 && tid < 64 + GBK) Xb[S][tid - 64] = xv[S];
    };
    // Tile t (of nt): multiply LDS buffer t % 2 while tile t + 2 loads into
    // register set t % 2, then move tile t + 1 from set (t + 1) % 2 into its
    // LDS buffer.
    auto step = [&](auto par, int t, int nt) {
        constexpr int P = decltype(par)::value;
        using Other = std::integral_constant<int, P ^ 1>;
        if (t + 2 < nt) load(par, (t + 2) * GBK);
#pragma unroll
        for (int s4 = 0; s4 < KS; s4 += 4) {   // four steps per 16-byte LDS read
            T a[2][4], b[GTJ][4];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int s = 0; s < 4; ++s) a[ti][s] = As[P][wr + 16 * ti + li][KS * lk + s4 + s];
#pragma unroll
            for (int tj = 0; tj < GTJ; ++tj)
#pragma unroll
                for (int s = 0; s < 4; ++s) b[tj][s] = Bt[P][wc + 16 * tj + li][KS * lk + s4 + s];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                    for (int tj = 0; tj < GTJ; ++tj)
                        acc[ti][tj] = MM::mma(a[ti][s], b[tj][s], acc[ti][tj]);
        }
        if (xrow && lane < GBN) {
#pragma unroll
            for (int s = 0; s < KS; ++s) er += Xa[P][KS * wave + s] * Bt[P][lane][KS * wave + s];
        }
        if (xcol) {
#pragma unroll
            for (int s = 0; s < KS; ++s) ec += As[P][lane][KS * wave + s] * Xb[P][KS * wave + s];
        }
        if (xrow && xcol && tid == 0) {
            for (int k = 0; k < GBK; ++k) ex += Xa[P][k] * Xb[P][k];
        }
        if (t + 1 < nt) store(Other{});
        __syncthreads();
    };
    using Zero = std::integral_constant<int, 0>;
    using One = std::integral_constant<int, 1>;
    const int nt = (M + GBK - 1) / GBK;   // k tiles (M >= 1)
    load(Zero{}, 0);
    store(Zero{});
    if (nt > 1) load(One{}, GBK);
    __syncthreads();
    for (int t = 0; t < nt; t += 2) {
        step(Zero{}, t, nt);
        if (t + 1 < nt) step(One{}, t + 1, nt);
    }

#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < GTJ; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = row0 + wr + 16 * ti + MM::row(lk, r);
                const int gc = col0 + wc + 16 * tj + li;
                if (gr < N && gc < N) {
                    // The k = M term.
                    T v = acc[ti][tj][r] + A[(size_t)gr * N + M] * B[(size_t)M * N + gc];
                    if (SCALE) v = (T)((double)v * solver_factor(f, gr, gc, N, sg));
                    C[(size_t)gr * N + gc] = v;
                }
            }
    if (xrow || xcol) {   // block-uniform
        red[0][wave][lane] = er;
        red[1][wave][lane] = ec;
        __syncthreads();
        if (tid < GBM) {
            const T amm = A[(size_t)M * N + M];
            if (xrow && tid < GBN) {
                const int gc = col0 + tid;
                T v = ((red[0][0][tid] + red[0][1][tid]) + (red[0][2][tid] + red[0][3][tid])) +
                      amm * B[(size_t)M * N + gc];
                if (SCALE) v = (T)((double)v * solver_factor(f, M, gc, N, sg));
                C[(size_t)M * N + gc] = v;
            }
            if (xcol && row0 + tid < N) {   // the row tile may extend past N
                const int gr = row0 + tid;
                T v = ((red[1][0][tid] + red[1][1][tid]) + (red[1][2][tid] + red[1][3][tid])) +
                      A[(size_t)gr * N + M] * B[(size_t)M * N + M];
                if (SCALE) v = (T)((double)v * solver_factor(f, gr, M, N, sg));
                C[(size_t)gr * N + M] = v;
            }
            if (xrow && xcol && tid == 0) {
                T v = ex + amm * B[(size_t)M * N + M];
                if (SCALE) v = (T)((double)v * solver_factor(f, M, M, N, sg));
                C[(size_t)M * N + M] = v;
            }
        }
    }
}

template <typename T, bool SCALE>
void gemm_launch(const GemmBatch<T> &bt, int N, const ScaleGeom &sg, hipStream_t s)
{
    static_assert(GemmTile<T>::BK <= 64, "the edge loaders use threads 0..127");
    static_assert(GBN % 32 == 0 && GBN <= 64 && GBM == 64, "wave layout 2 x 2");
    const dim3 grid((N - 1 + GBN - 1) / GBN, (N - 1 + GBM - 1) / GBM, 7);
    k_gemm_mfma<T, SCALE><<<grid, 256, 0, s>>>(bt, N, sg);
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
    k_rhs<TG, TPRED><<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(rhs, w1);
    hipError_t e;
    if (timer && (e = timer->begin(KC_SOLVE, 4, s)) != hipSuccess) return e;
    gemm_launch<TG, false>(fy, N, sg, s);
    gemm_launch<TG, true>(sx, N, sg, s);
    gemm_launch<TG, false>(ix, N, sg, s);
    gemm_launch<TG, false>(iy, N, sg, s);
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
