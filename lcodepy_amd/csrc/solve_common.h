// Pieces shared by the two implementations of the field solve
// (plasma3d/fields.py:112-322): field order, transform kinds, the right-hand
// sides and the output assembly.
#pragma once
#include "common.h"
#include "kernels.h"

namespace lc3d {

// Field order of the solve.
enum Field { F_EX = 0, F_EY, F_BX, F_BY, F_EZ, F_BZ, F_PHI };

// Transform along y (the contiguous index j) for each field: DST-I for the
// fields that vanish at the y edges, DCT-I otherwise; along x likewise
// (fields.py:140-266).
__host__ __device__ __forceinline__ bool dst_along_y(int f) { return f == F_EX || f == F_BY || f == F_EZ || f == F_PHI; }
__host__ __device__ __forceinline__ bool dst_along_x(int f) { return f == F_EY || f == F_BX || f == F_EZ || f == F_PHI; }

// Grid points whose right-hand side enters the transform: the DST directions
// skip their two edge points.
__device__ __forceinline__ bool in_rhs_domain(int f, int i, int j, int N) {
    return (!dst_along_x(f) || (i > 0 && i < N - 1)) && (!dst_along_y(f) || (j > 0 && j < N - 1));
}

struct ScaleGeom {
    const double *lam;
    double c;      // subtraction coefficient of the mixed fields
    double d2;     // (2 (N - 1))^2
};

// The solver-matrix factor of field f at spectral indices (kx, ky) = (gr, gc)
// (initialization.py:15-67): zero where a DST index is an edge and at [0, 0]
// for the Neumann field Bz.
__device__ __forceinline__ double solver_factor(int f, int gr, int gc, int N, const ScaleGeom &sg)
{
    const int M = N - 1;
    const bool dx = dst_along_x(f), dy = dst_along_y(f);
    const double cc = (f == F_EZ || f == F_BZ || f == F_PHI) ? 0.0 : sg.c;
    const bool ok = (!dx || (gr > 0 && gr < M)) && (!dy || (gc > 0 && gc < M)) &&
                    !(f == F_BZ && gr == 0 && gc == 0);
    return ok ? (1.0 / (sg.lam[gr] + sg.lam[gc] + cc)) / sg.d2 : 0.0;
}

// Right-hand side of field f at (i, j) from the currents of layers k (P) and
// k+1 (F), the beam densities and the predictor fields (fields.py:112-266).
// Valid on in_rhs_domain(); the x/y stencils stay inside the grid there.
template <typename TG, typename TPRED> struct RhsEval {
    SolveIn<TG, TPRED> in;
    int N;
    TG inv2h, invdxi, c;
    __device__ TG roF(int o) const { return in.roF[0][o] + in.roF[1][o]; }
    __device__ TG jxF(int o) const { return in.jxF[0][o] + in.jxF[1][o]; }
    __device__ TG jyF(int o) const { return in.jyF[0][o] + in.jyF[1][o]; }
    __device__ TG jzF(int o) const { return in.jzF[0][o] + in.jzF[1][o]; }
    __device__ TG roP(int o) const { return in.roP[0][o] + in.roP[1][o]; }
    __device__ TG jxP(int o) const { return in.jxP[0][o] + in.jxP[1][o]; }
    __device__ TG jyP(int o) const { return in.jyP[0][o] + in.jyP[1][o]; }
    __device__ TG jzP(int o) const { return in.jzP[0][o] + in.jzP[1][o]; }
    __device__ TG ro_half(int o) const {
        return (roF(o) + (TG)in.rbF[o] + roP(o) + (TG)in.rbP[o]) * TG(0.5);
    }
    __device__ TG jz_half(int o) const {
        return (jzF(o) + (TG)in.rbF[o] + jzP(o) + (TG)in.rbP[o]) * TG(0.5);
    }

    __device__ TG operator()(int f, int i, int j) const {
        const int o = i * N + j;
        const bool inner = i > 0 && i < N - 1 && j > 0 && j < N - 1;
        switch (f) {
        case F_EX: {
            const TG d = inner ? (ro_half(o + N) - ro_half(o - N)) * inv2h : TG(0);
            return -((d - (jxP(o) - jxF(o)) * invdxi) - c * (TG)in.ExP[o]);
        }
        case F_EY: {
            const TG d = inner ? (ro_half(o + 1) - ro_half(o - 1)) * inv2h : TG(0);
            return -((d - (jyP(o) - jyF(o)) * invdxi) - c * (TG)in.EyP[o]);
        }
        case F_BX: {
            const TG d = inner ? (jz_half(o + 1) - jz_half(o - 1)) * inv2h : TG(0);
            return (d - (jyP(o) - jyF(o)) * invdxi) + c * (TG)in.BxP[o];
        }
        case F_BY: {
            const TG d = inner ? (jz_half(o + N) - jz_half(o - N)) * inv2h : TG(0);
            return -((d - (jxP(o) - jxF(o)) * invdxi) - c * (TG)in.ByP[o]);
        }
        case F_EZ:
            return -((jxF(o + N) - jxF(o - N)) + (jyF(o + 1) - jyF(o - 1))) * inv2h;
        case F_BZ: {
            const TG a = inner ? jxF(o + 1) - jxF(o - 1) : TG(0);
            const TG b = inner ? jyF(o + N) - jyF(o - N) : TG(0);
            return -(a - b) * inv2h;
        }
        default:   // F_PHI
            return roF(o) - jzF(o);
        }
    }
};

// Writes the solved value r of field f at flat index o into the outputs
// (fields.py:299-314): full = 2 half - prev and half for Ex Ey Bx By,
// full and half = (full + prev)/2 for Ez, the raw Bz (its mean is removed
// afterwards), Phi into both.
template <typename TG, typename TFULL, typename THALF>
__device__ __forceinline__ void emit_output(const SolveOut<TFULL, THALF> &out, int f, size_t o,
                                            TG r, bool write_full)
{
    switch (f) {
    case F_EX: case F_EY: case F_BX: case F_BY:
        out.half[f][o] = (THALF)r;
        if (write_full) out.full[f][o] = (TFULL)(2 * (TFULL)r - out.prev[f][o]);
        break;
    case F_EZ: {
        const TFULL v = (TFULL)r;
        if (write_full) out.full[f][o] = v;
        out.half[f][o] = (THALF)((v + out.prev[f][o]) / 2);
        break;
    }
    case F_PHI:
        out.full[f][o] = (TFULL)r;   // fields_half.Phi is fields_full.Phi
        out.half[f][o] = (THALF)r;
        break;
    default:                         // F_BZ
        out.full[f][o] = (TFULL)r;
        break;
    }
}

// Subtracts the mean of Bz (rowsum holds its N row sums) and forms the
// half-step Bz (fields.py:264, 312).
template <typename TFULL, typename THALF>
__global__ __launch_bounds__(256) void k_bz_finalize(int N, const double *__restrict__ rowsum,
                                                     const TFULL *__restrict__ prev,
                                                     TFULL *__restrict__ full,
                                                     THALF *__restrict__ half)
{
    __shared__ double red[256];
    double s = 0;
    for (int i = threadIdx.x; i < N; i += 256) s += rowsum[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const double mean = red[0] / ((double)N * N);
    const int64_t n = (int64_t)N * N;
    for (int64_t o = blockIdx.x * (int64_t)256 + threadIdx.x; o < n; o += (int64_t)gridDim.x * 256) {
        const TFULL v = (TFULL)((double)full[o] - mean);
        full[o] = v;
        half[o] = (THALF)((v + prev[o]) / 2);
    }
}

}  // namespace lc3d
