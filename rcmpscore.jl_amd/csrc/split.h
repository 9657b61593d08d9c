// Split-precision tile products for the trajectory kernel: FP64 panels are cut into exact
// 8-bit integer digits and multiplied on the INT8 MFMA with exact INT32 accumulation.
//
// A panel P (16 x 64 row block, or 64 x 16 column block stored transposed) with largest entry
// magnitude m gets a power-of-two scale S with m 2^S in [2^45, 2^46).  Every entry x becomes the
// 48-bit integer X = x 2^S (rounded once, |X| < 2^46) written in six balanced digits,
//     X = sum_d s_d 2^(8 (5 - d)),  s_d in [-128, 127],
// so a product of two panels is sum_{d,e} 2^(8 (10 - d - e) - S_A - S_B) (s^A_d . s^B_e).  Each
// digit product over k = 64 is exact in INT32 (|.| <= 64 * 2^14), as is the sum of the up-to-seven
// products of one anti-diagonal c = d + e; the diagonals c <= RC_SPLIT_DIAG - 1 are kept and
// combined in FP64.  The dropped diagonals are below 2^-47 of |A| |B| k-sums, the rounding of
// X below 2^-46 of each panel's maximum: the product error is about 2^-45 relative to
// max|A| max|B| k, against 2^-52 for the FP64 MFMA.
//
// Digit planes live in LDS as [digit][part re/im][16 rows or columns][k] bytes, k contiguous
// with a row stride of 80 bytes (the 16 lanes of one MFMA operand read hit disjoint banks).
#pragma once
#include "device_common.h"

typedef int v4i __attribute__((ext_vector_type(4)));

#define RC_SPLIT_DIGITS 6
#ifndef RC_SPLIT_DIAG
#define RC_SPLIT_DIAG 7
#endif
#define RC_PLANE_STRIDE 80
#define RC_PLANE_BYTES (16 * RC_PLANE_STRIDE)
#define RC_PANEL_BYTES (RC_SPLIT_DIGITS * 2 * RC_PLANE_BYTES)

// scale exponent for a panel whose largest magnitude is m (0 for an all-zero panel)
__host__ __device__ __forceinline__ int rc_split_scale(double m) {
    return m > 0.0 ? 45 - ilogb(m) : 0;
}

// x 2^S (|x 2^S| < 2^46) -> digits s_0..s_2 in bytes 2,1,0 of hd and s_3..s_5 in bytes 2,1,0 of ld
__host__ __device__ __forceinline__ void rc_digits(double x, int S, uint32_t& hd, uint32_t& ld) {
    const double xs = ldexp(x, S - 24);          // |xs| < 2^22
    int hi = (int)xs;                             // truncation; the remainder keeps the sign of xs
    const int lo = (int)rint(ldexp(xs - (double)hi, 24));   // |lo| <= 2^24
    const int lo2 = lo + 0x808080;
    hi += lo2 >> 24;                              // carry of the balanced low digits, in {-1, 0, 1}
    const int hi2 = hi + 0x808080;                // in (0, 2^24): balanced high digits, no carry out
    hd = ((uint32_t)hi2 & 0xFFFFFFu) ^ 0x808080u;
    ld = ((uint32_t)lo2 & 0xFFFFFFu) ^ 0x808080u;
}

// byte q (0..3) of digit d of an element from its packed (hd, ld)
__host__ __device__ __forceinline__ uint32_t rc_digit_byte(uint32_t hd, uint32_t ld, int d) {
    const uint32_t w = d < 3 ? hd : ld;
    const int sh = 8 * (2 - (d < 3 ? d : d - 3));
    return (w >> sh) & 0xFFu;
}

// Write four consecutive-k elements (re and im) of row/column `i` at k = k0..k0+3 into the
// panel's digit planes (one 32-bit LDS store per digit and part).
__device__ __forceinline__ void rc_store_digits4(unsigned char* panel, int i, int k0, const cplx* v, int S) {
    uint32_t hr[4], lr[4], hm[4], lm[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        rc_digits(v[q].re, S, hr[q], lr[q]);
        rc_digits(v[q].im, S, hm[q], lm[q]);
    }
#pragma unroll
    for (int d = 0; d < RC_SPLIT_DIGITS; ++d) {
        uint32_t wr = 0, wi = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            wr |= rc_digit_byte(hr[q], lr[q], d) << (8 * q);
            wi |= rc_digit_byte(hm[q], lm[q], d) << (8 * q);
        }
        unsigned char* base = panel + (size_t)(d * 2) * RC_PLANE_BYTES + i * RC_PLANE_STRIDE + k0;
        *reinterpret_cast<uint32_t*>(base) = wr;
        *reinterpret_cast<uint32_t*>(base + RC_PLANE_BYTES) = wi;
    }
}

// MFMA operand: lane l reads row/column l % 16, k = 16 (l / 16) .. + 15 of digit d, part p.
__device__ __forceinline__ v4i rc_operand(const unsigned char* panel, int d, int p, int lane) {
    return *reinterpret_cast<const v4i*>(panel + (size_t)(d * 2 + p) * RC_PLANE_BYTES +
                                          (lane & 15) * RC_PLANE_STRIDE + 16 * (lane >> 4));
}

// This wave's real product (part pa of A times part pb of B) added in FP64 to out[4]; out[q] is
// element (4 (lane/16) + q, lane%16) of the 16 x 16 tile (the INT32 MFMA result layout; the FP64
// MFMA's is (4q + lane/16, lane%16)).  scale2 = -(S_A + S_B).
__device__ __forceinline__ void rc_split_product(double* out, const unsigned char* A, int pa,
                                                 const unsigned char* B, int pb, int scale2, int lane) {
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
    // sum_c acc_c 2^(8 (10 - c)) 2^scale2, Horner from the leading diagonal
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = (double)acc[0][q];
#pragma unroll
        for (int c = 1; c < RC_SPLIT_DIAG; ++c) v = fma(v, 256.0, (double)acc[c][q]);
        out[q] += ldexp(v, scale2 + 8 * (10 - (RC_SPLIT_DIAG - 1)));
    }
}
