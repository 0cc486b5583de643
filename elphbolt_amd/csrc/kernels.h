// Device kernels of the vertex path (gfx950).  Layout conventions:
//   complex128 values are double2 (x = re, y = im);
//   matrices are column-major unless a comment says otherwise.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace elph {

typedef double __attribute__((ext_vector_type(4))) d4;
// native 2-double vector for staging copies (HIP's double2 is a union-based
// struct that the compiler does not keep in registers when arrays of it are
// carried across branches)
typedef double __attribute__((ext_vector_type(2))) dv2;

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

typedef int v4i __attribute__((ext_vector_type(4)));

// D(16x16, int32) += A(16x64, int8) B(64x16, int8), exact.  Operand layout
// (checked on gfx950 by a layout probe): lane i + 16 g holds A[i][16 g .. 16 g + 15]
// (16 bytes) and B[16 g .. 16 g + 15][i]; D[i][j] is element i % 4 of lane
// j + 16 (i / 4).
__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Digits per part in the mixed-precision products (base 254, see split_digits).
constexpr int kNS = 4;

// Signed-digit split of x in [-1, 1] into NS int8 digits d_s in [-127, 127]:
// x ~= sum_s d_s / (127 * 254^s), error below 1 / (2 * 127 * 254^(NS-1)).
// The remainder is carried scaled to the current digit (y_0 = 127 x, y_{s+1}
// = 254 (y_s - d_s), |y_s - d_s| <= 1/2 exactly, so |d_s| <= 127), which
// needs no division.
template <int NS>
__host__ __device__ inline void split_digits(double x, signed char (&d)[NS]) {
    double y = x * 127.0;
    for (int s = 0; s < NS; ++s) {
        const double q = rint(y);
        d[s] = (signed char)q;
        y = (y - q) * 254.0;
    }
}

// Packs 16 signed bytes (one per b of an unrolled loop) into a v4i: byte b of
// word b / 4 (little endian), as the int8 MFMA operand registers hold them.
__device__ __forceinline__ void put_byte(unsigned (&w)[4], int b, signed char v) {
    w[b >> 2] |= (unsigned)(unsigned char)v << (8 * (b & 3));
}
__device__ __forceinline__ v4i as_v4i(const unsigned (&w)[4]) {
    return v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

// Phase-table index (c . R) mod L of byte-packed coefficients c (c_d < L <= 255)
// and cell R (R_d < 256): one 4-way byte dot product, x <= 3 * 255 * 255 < 2^18,
// then the quotient floor(x / L) as the high word of x * ceil(2^32 / L), which
// is exact for x < 2^32 / (L - 1) > 2^24.
__host__ __device__ inline unsigned phase_lmul(int64_t L) {
    return L <= 1 ? 0u : (unsigned)(((1ull << 32) + (uint64_t)L - 1) / (uint64_t)L);
}
__device__ __forceinline__ int phase_index(unsigned cpk, unsigned rp, unsigned lmul, int negL) {
    const unsigned x = __builtin_amdgcn_udot4(cpk, rp, 0u, false);
    return __mul24((int)__umulhi(x, lmul), negL) + (int)x;
}

// Asynchronous copy global -> LDS of one 1 KiB row per wave (16 bytes per lane,
// lane i to dst + 16 i); dst must be the same for all lanes of the wave.  The
// data is visible to the wave after lds_dma_wait() and to the block after a
// following barrier.
__device__ __forceinline__ void lds_dma16(const void *src, void *dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}
// Same with 4 bytes per lane (lane i to dst + 4 i).
__device__ __forceinline__ void lds_dma4(const void *src, void *dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                     (__attribute__((address_space(3))) void *)dst, 4, 0, 0);
}
__device__ __forceinline__ void lds_dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// GEMM core on the FP64 matrix cores, one 256-thread block (4 waves):
//   (cr + i ci)[rb][c] (16 rows x 16 cols) += sum_K A(rb, K) * B[K][c]
// for the wave's RB row blocks (A supplied by aop(rb, K), K the global row
// index of B, evaluated by lane (li, lk) for its row li of block rb and its K
// slot) and the CB column blocks of the tile.  B (Kp rows of 16*CB complex,
// row stride ldb) is streamed through LDS in KC-row chunks (buf0/buf1, KC*16*CB
// complex each) with a register prefetch of the next chunk.  Kp % KC == 0.
// aop returns double2 (complex A: 4 real products per K) or double (real A:
// 2 real products per K).
template <int CB, int RB, int KC, typename AOp>
__device__ __forceinline__ void gemm_core_f64(const double2 *__restrict__ B, int ldb, int Kp, AOp aop,
                                              double2 *buf0, double2 *buf1, d4 (&cr)[RB][CB], d4 (&ci)[RB][CB]) {
    constexpr int CP = 16 * CB;
    const int tid = threadIdx.x, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < CB; ++c) { cr[rb][c] = d4{0, 0, 0, 0}; ci[rb][c] = d4{0, 0, 0, 0}; }
    constexpr int NPF = (KC * CP + 255) / 256;
    constexpr int chunk = KC * CP;
    // with chunk a multiple of the block size every prefetch slot is live; the
    // guard would make the compiler keep pf[] in scratch memory
    constexpr bool exact = chunk % 256 == 0;
    for (int i = tid; i < chunk; i += 256) buf0[i] = B[(size_t)(i / CP) * ldb + i % CP];
    __syncthreads();
    const int nchunks = Kp / KC;
    for (int ch = 0; ch < nchunks; ++ch) {
        dv2 pf[NPF];
        const bool more = ch + 1 < nchunks;
        if (more) {
            const dv2 *src = reinterpret_cast<const dv2 *>(B + (size_t)(ch + 1) * KC * ldb);
#pragma unroll
            for (int j = 0; j < NPF; ++j) {
                const int i = tid + 256 * j;
                if (exact || i < chunk) pf[j] = src[(size_t)(i / CP) * ldb + i % CP];
            }
        }
        const double2 *Bs = (ch & 1) ? buf1 : buf0;
#pragma unroll 2
        for (int ks = 0; ks < KC / 4; ++ks) {
            const int kl = ks * 4 + lk;
            if constexpr (std::is_same<decltype(aop(0, 0)), double>::value) {
                double av[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) av[rb] = aop(rb, ch * KC + kl);
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const double2 b = Bs[kl * CP + c * 16 + li];
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) {
                        cr[rb][c] = mfma_f64(av[rb], b.x, cr[rb][c]);
                        ci[rb][c] = mfma_f64(av[rb], b.y, ci[rb][c]);
                    }
                }
            } else {
                double ar[RB], ai[RB], nai[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const double2 t = aop(rb, ch * KC + kl);
                    ar[rb] = t.x;
                    ai[rb] = t.y;
                    nai[rb] = -t.y;
                }
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const double2 b = Bs[kl * CP + c * 16 + li];
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) cmfma(ar[rb], ai[rb], nai[rb], b.x, b.y, cr[rb][c], ci[rb][c]);
                }
            }
        }
        if (more) {
            dv2 *dst = reinterpret_cast<dv2 *>((ch & 1) ? buf0 : buf1);
#pragma unroll
            for (int j = 0; j < NPF; ++j) {
                const int i = tid + 256 * j;
                if (exact || i < chunk) dst[i] = pf[j];
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int pos_mod(int x, int L) {
    int m = x % L;
    return m < 0 ? m + L : m;
}

}  // namespace elph
