// Device kernels of the vertex path (gfx950).  Layout conventions:
//   complex128 values are double2 (x = re, y = im);
//   matrices are column-major unless a comment says otherwise.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace elph {

typedef double __attribute__((ext_vector_type(4))) d4;

// D(16x16) += A(16x4) B(4x16) in FP64 on the matrix cores.  Operand layout
// (checked by elph_selftest_mfma): A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4 of the vector.
__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Complex 16x16x4 MAC: (cr, ci) += (ar + i ai)(br + i bi); nai = -ai.
__device__ __forceinline__ void cmfma(double ar, double ai, double nai, double br, double bi, d4 &cr, d4 &ci) {
    cr = mfma_f64(ar, br, cr);
    cr = mfma_f64(nai, bi, cr);
    ci = mfma_f64(ar, bi, ci);
    ci = mfma_f64(ai, br, ci);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}

__device__ __forceinline__ int pos_mod(int x, int L) {
    int m = x % L;
    return m < 0 ? m + L : m;
}

}  // namespace elph
