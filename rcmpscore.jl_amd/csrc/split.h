// Split-precision tile products for the trajectory kernel: FP64 panels are cut into exact
// 8-bit integer digits and multiplied on the INT8 MFMA with exact INT32 accumulation.
//
// Each row of a row panel (16 x 64 row block), each column of a column panel (64 x 16 column
// block stored transposed), or a whole constant panel, with largest entry magnitude m gets a
// power-of-two scale S with m 2^S in [2^45, 2^46).  Every entry x becomes the 48-bit integer
// X = x 2^S (rounded once, |X| < 2^46) written in six balanced digits (the default cut),
//     X = sum_d s_d 2^(8 (5 - d)),  s_d in [-128, 127],
// so a product of two panels is sum_{d,e} 2^(8 (10 - d - e) - S_A - S_B) (s^A_d . s^B_e).  Each
// digit product over k = 64 is exact in INT32 (|.| <= 64 * 2^14), as is the sum of the up-to-seven
// products of one anti-diagonal c = d + e; the diagonals c <= RC_SPLIT_DIAG - 1 are kept and
// combined in FP64.  The dropped diagonals are below 2^-47 of |A| |B| k-sums, the rounding of
// X below 2^-46 of each row's (column's) maximum: the error of an output element is about 2^-45
// relative to max|A row| max|B column| k, against 2^-52 for the FP64 MFMA.
//
// RC_SPLIT_DIGITS (n = 4..6) and RC_SPLIT_DIAG select other cuts (integers of 8n - 2 bits, n (n+1)/2
// digit products per real product at RC_SPLIT_DIAG = n).  Six digits are the default because the
// step controllers need them: on the D = 64 workload 5 digits (error about 2^-37) leave the energy
// within 3e-12 of FP64 but the adjoint systems' rejections multiply the attempts 4-6 fold, and
// 4 digits by several hundred, since the product error then exceeds atol on small entries of a
// row whose maximum sets the scale.
//
// Digit planes live in LDS as [digit][part re/im][16 rows or columns][k] bytes, k contiguous
// with a row stride of 80 bytes (the 16 lanes of one MFMA operand read hit disjoint banks).
#pragma once
#include "device_common.h"

typedef int v4i __attribute__((ext_vector_type(4)));

#ifndef RC_SPLIT_DIGITS
#define RC_SPLIT_DIGITS 6
#endif
#ifndef RC_SPLIT_DIAG
#define RC_SPLIT_DIAG (RC_SPLIT_DIGITS + 1)
#endif
static_assert(RC_SPLIT_DIGITS >= 4 && RC_SPLIT_DIGITS <= 6, "4 to 6 digits per operand");
static_assert(RC_SPLIT_DIAG >= 1 && RC_SPLIT_DIAG <= 2 * RC_SPLIT_DIGITS - 1, "kept anti-diagonals");
// integers carry RC_SPLIT_BITS = 8 n - 2 bits: |X| < 2^(8 n - 2)
#define RC_SPLIT_TOP (8 * RC_SPLIT_DIGITS - 3)
#define RC_PLANE_STRIDE 80
#define RC_PLANE_BYTES (16 * RC_PLANE_STRIDE)
#define RC_PANEL_BYTES (RC_SPLIT_DIGITS * 2 * RC_PLANE_BYTES)

// scale exponent for a panel whose largest magnitude is m (0 for an all-zero panel), at most
// RC_SPLIT_SMAX so that 2^S stays finite (entries below 2^-(RC_SPLIT_SMAX - 45) then become zero)
#define RC_SPLIT_SMAX 1000
__host__ __device__ __forceinline__ int rc_split_scale(double m) {
    return m > 0.0 ? min(RC_SPLIT_TOP - ilogb(m), RC_SPLIT_SMAX) : 0;
}
// the same from the biased FP64 exponent e of m (e = 0: zero or subnormal, treated as zero)
__host__ __device__ __forceinline__ int rc_split_scale_e(int e) {
    return e > 0 ? min(RC_SPLIT_TOP - (e - 1023), RC_SPLIT_SMAX) : 0;
}

// B = 0x80..80 (128 in each of the n digit bytes): the bias that makes balanced digits bytes
#define RC_SPLIT_BIAS_HI (RC_SPLIT_DIGITS == 6 ? 0x8080u : RC_SPLIT_DIGITS == 5 ? 0x80u : 0u)
#define RC_SPLIT_MAGIC (4503599627370496.0 + 2155905152.0 + 4294967296.0 * (double)RC_SPLIT_BIAS_HI)

// x 2^S -> the (8n - 2)-bit integer X = rint(x 2^S) in n balanced digits.  fma(x, 2^S, 2^52 + B)
// rounds once to the integer 2^52 + B + X (|X| < 2^(8n - 2), so the sum stays in [2^52, 2^53) where
// the ulp is 1); its low 8n mantissa bits are Y = X + B, and byte b of Y XOR 0x80 is the digit of
// weight 256^b.  lo holds bytes 0-3, hi bytes 4 and 5 (as many as there are digits).
__host__ __device__ __forceinline__ void rc_digits(double x, double scale, uint32_t& lo, uint32_t& hi) {
    const uint64_t u = __builtin_bit_cast(uint64_t, fma(x, scale, RC_SPLIT_MAGIC));
    lo = (uint32_t)u ^ 0x80808080u;
    hi = (uint32_t)(u >> 32) ^ RC_SPLIT_BIAS_HI;
}

// byte d (0..3) of each of w0..w3 gathered into out[d] (v_perm_b32: selector bytes 0-3 address the
// second operand, 4-7 the first)
__device__ __forceinline__ void rc_bytes_t4(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t* out) {
    const uint32_t a_lo = __builtin_amdgcn_perm(w1, w0, 0x05010400u), a_hi = __builtin_amdgcn_perm(w1, w0, 0x07030602u);
    const uint32_t b_lo = __builtin_amdgcn_perm(w3, w2, 0x05010400u), b_hi = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
    out[0] = __builtin_amdgcn_perm(b_lo, a_lo, 0x05040100u);
    out[1] = __builtin_amdgcn_perm(b_lo, a_lo, 0x07060302u);
    out[2] = __builtin_amdgcn_perm(b_hi, a_hi, 0x05040100u);
    out[3] = __builtin_amdgcn_perm(b_hi, a_hi, 0x07060302u);
}
__device__ __forceinline__ void rc_bytes_t2(uint32_t h0, uint32_t h1, uint32_t h2, uint32_t h3, uint32_t* out) {
    const uint32_t a = __builtin_amdgcn_perm(h1, h0, 0x05010400u), b = __builtin_amdgcn_perm(h3, h2, 0x05010400u);
    out[0] = __builtin_amdgcn_perm(b, a, 0x05040100u);
    out[1] = __builtin_amdgcn_perm(b, a, 0x07060302u);
}

// Write four consecutive-k elements (re and im) of row/column `i` at k = k0..k0+3 into the
// panel's digit planes: one 32-bit LDS store per digit and part.  scale = 2^S.
__device__ __forceinline__ void rc_store_digits4(unsigned char* panel, int i, int k0, const cplx* v, double scale) {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        uint32_t lo[4], hi[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) rc_digits(part ? v[q].im : v[q].re, scale, lo[q], hi[q]);
        uint32_t bytes[6];
        rc_bytes_t4(lo[0], lo[1], lo[2], lo[3], bytes);            // bytes[b]: digit of weight 256^b
        if constexpr (RC_SPLIT_DIGITS > 4) rc_bytes_t2(hi[0], hi[1], hi[2], hi[3], bytes + 4);
        unsigned char* base = panel + (size_t)part * RC_PLANE_BYTES + i * RC_PLANE_STRIDE + k0;
        // plane d holds the digit of weight 256^(n-1-d)
#pragma unroll
        for (int d = 0; d < RC_SPLIT_DIGITS; ++d)
            *reinterpret_cast<uint32_t*>(base + d * 2 * RC_PLANE_BYTES) = bytes[RC_SPLIT_DIGITS - 1 - d];
    }
}

// MFMA operand: lane l reads row/column l % 16, k = 16 (l / 16) .. + 15 of digit d, part p.
__device__ __forceinline__ v4i rc_operand(const unsigned char* panel, int d, int p, int lane) {
    return *reinterpret_cast<const v4i*>(panel + (size_t)(d * 2 + p) * RC_PLANE_BYTES +
                                          (lane & 15) * RC_PLANE_STRIDE + 16 * (lane >> 4));
}

// This wave's real product (part pa of A times part pb of B) added in FP64 to out[4]; out[q] is
// element (4 (lane/16) + q, lane%16) of the 16 x 16 tile (the INT32 MFMA result layout; the FP64
// MFMA's is (4q + lane/16, lane%16)).  scale2[q] = -(S_A(row) + S_B(column)) of that element: A may
// carry one scale per row and B one per column, since each output element's k-sum uses a single
// row of A and a single column of B.
__device__ __forceinline__ void rc_split_product(double* out, const unsigned char* A, int pa,
                                                 const unsigned char* B, int pb, const int* scale2, int lane) {
    v4i a[RC_SPLIT_DIGITS];
#pragma unroll
    for (int d = 0; d < RC_SPLIT_DIGITS; ++d) a[d] = rc_operand(A, d, pa, lane);
    v4i acc[RC_SPLIT_DIAG];
#pragma unroll
    for (int c = 0; c < RC_SPLIT_DIAG; ++c) acc[c] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < RC_SPLIT_DIGITS; ++e) {
        const v4i be = rc_operand(B, e, pb, lane);
#pragma unroll
        for (int d = 0; d < RC_SPLIT_DIGITS; ++d)
            if (d + e < RC_SPLIT_DIAG) acc[d + e] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[d], be, acc[d + e], 0, 0, 0);
    }
    // sum_c acc_c 2^(8 (2 (n - 1) - c)) 2^scale2, Horner from the leading diagonal
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = (double)acc[0][q];
#pragma unroll
        for (int c = 1; c < RC_SPLIT_DIAG; ++c) v = fma(v, 256.0, (double)acc[c][q]);
        out[q] += ldexp(v, scale2[q] + 8 * (2 * (RC_SPLIT_DIGITS - 1) - (RC_SPLIT_DIAG - 1)));
    }
}
