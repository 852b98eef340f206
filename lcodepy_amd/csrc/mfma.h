// 16x16x4 matrix-core products (CDNA3/CDNA4 v_mfma_f32_16x16x4_f32 and
// v_mfma_f64_16x16x4_f64), one wave:
//   D[16][16] += A[16][4] B[4][16]
// Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].  Element r of its
// accumulator holds D[row(l / 16, r)][l % 16], where row(q, r) is 4 q + r for
// the float32 instruction and q + 4 r for the float64 one (measured on gfx950
// with operands that encode their indices).
#pragma once
#include <hip/hip_runtime.h>

namespace lc3d {

typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef double v4d_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma16;

template <> struct Mfma16<float> {
    using acc_t = v4f_t;
    static __device__ __forceinline__ int row(int q, int r) { return 4 * q + r; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <> struct Mfma16<double> {
    using acc_t = v4d_t;
    static __device__ __forceinline__ int row(int q, int r) { return q + 4 * r; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

}  // namespace lc3d
