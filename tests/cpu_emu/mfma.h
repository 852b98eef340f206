// CPU emulation of lcodepy_amd/csrc/mfma.h: the 16x16x4 wave-level product
// with the operand and result layout the matrix cores use (lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; element r of its result is
// D[row(l / 16, r)][l % 16], row(q, r) = 4 q + r for float32, q + 4 r for
// float64).
// The 64 lanes of a wave exchange their operands through per-wave arrays
// between two wave barriers.
#pragma once
#include "hip/hip_runtime.h"

namespace lc3d {

template <typename T> struct Acc4 {
    T v[4];
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
};

template <typename T> struct MfmaLds {
    static T &a(int w, int i, int k) { thread_local T s[16][16][4]; return s[w][i][k]; }
    static T &b(int w, int k, int j) { thread_local T s[16][4][16]; return s[w][k][j]; }
};

template <typename T> struct Mfma16 {
    using acc_t = Acc4<T>;
    static int row(int q, int r) { return sizeof(T) == 4 ? 4 * q + r : q + 4 * r; }
    static acc_t mma(T a, T b, acc_t c) {
        const int t = threadIdx.x, w = t / 64, l = t % 64;
        MfmaLds<T>::a(w, l % 16, l / 16) = a;
        MfmaLds<T>::b(w, l / 16, l % 16) = b;
        emu_sync_wave();
        acc_t d;
        for (int r = 0; r < 4; ++r) {
            T s = c[r];
            for (int k = 0; k < 4; ++k) s += MfmaLds<T>::a(w, row(l / 16, r), k) * MfmaLds<T>::b(w, k, l % 16);
            d[r] = s;
        }
        emu_sync_wave();
        return d;
    }
};

}  // namespace lc3d
