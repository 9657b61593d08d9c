// Device helpers shared by the HIP kernels: complex arithmetic, FP64 MFMA complex tiles,
// agent-coherent 16-byte loads/stores for data exchanged between workgroups, the source
// function table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmake(double r, double i) { cplx z; z.re = r; z.im = i; return z; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cmake(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cmake(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return cmake(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cplx cscale(double s, cplx a) { return cmake(s * a.re, s * a.im); }
__device__ __forceinline__ cplx cconj(cplx a) { return cmake(a.re, -a.im); }
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
// |z| as hypot, matching the host's cabs within an ulp
__device__ __forceinline__ double cabsd(cplx a) { return hypot(a.re, a.im); }

// ---------------------------------------------------------------------------------------
// Agent-coherent (sc1) accesses through buffer resources.  Stores write through the
// workgroup's L2, loads are served coherently, so data produced by a workgroup on another
// XCD is seen once its flag is seen.  num_records bounds every access: an access outside
// [0, bytes) reads zero / is dropped instead of faulting.
// ---------------------------------------------------------------------------------------
#define RC_COHERENT_POLICY 16  // CPol::SC1
#define RC_BUF_DWORD3 0x00020000

// wave-uniform copies of values the compiler cannot prove uniform (control state, addresses)
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double unid(double x) { return __builtin_bit_cast(double, uni64(__builtin_bit_cast(uint64_t, x))); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rc_rsrc(const void* base, uint32_t bytes) {
    void* b = reinterpret_cast<void*>(uni64(reinterpret_cast<uint64_t>(base)));
    return __builtin_amdgcn_make_buffer_rsrc(b, 0, (int)__builtin_amdgcn_readfirstlane(bytes), RC_BUF_DWORD3);
}

__device__ __forceinline__ cplx ld_coh(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_offset, 0, RC_COHERENT_POLICY);
    cplx z;
    z.re = __builtin_bit_cast(double, (uint64_t)v.x | ((uint64_t)v.y << 32));
    z.im = __builtin_bit_cast(double, (uint64_t)v.z | ((uint64_t)v.w << 32));
    return z;
}

__device__ __forceinline__ void st_coh(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset, cplx z) {
    uint64_t a = __builtin_bit_cast(uint64_t, z.re), b = __builtin_bit_cast(uint64_t, z.im);
    v4u v;
    v.x = (uint32_t)a; v.y = (uint32_t)(a >> 32); v.z = (uint32_t)b; v.w = (uint32_t)(b >> 32);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, byte_offset, 0, RC_COHERENT_POLICY);
}

__device__ __forceinline__ double ld_coh_f64(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh_f64(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// read-only data written before the launch, read at a wave-uniform index: through the scalar cache
__device__ __forceinline__ double ld_const(const double* p, int i) {
    return ((const __attribute__((address_space(4))) double*)p)[__builtin_amdgcn_readfirstlane(i)];
}
__device__ __forceinline__ uint32_t ld_flag(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_flag(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the 256 threads' (re, im) into out[0..1] by the pairwise tree t += t + s, s = 128, 64,
// ..., 1: the first two levels through LDS, the last six inside wave 0 by shuffles (lane t adds
// lane t + s's value to its own, the operands and order of the LDS tree).
__device__ __forceinline__ void tree_sum_256(double re, double im, double* sre, double* sim, double* out) {
    const int t = threadIdx.x;
    sre[t] = re;
    sim[t] = im;
    __syncthreads();
    if (t < 128) {
        sre[t] += sre[t + 128];
        sim[t] += sim[t + 128];
    }
    __syncthreads();
    if (t < 64) {
        re = sre[t] + sre[t + 64];
        im = sim[t] + sim[t + 64];
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) {
            re += __shfl_down(re, s);
            im += __shfl_down(im, s);
        }
        if (t == 0) {
            out[0] = re;
            out[1] = im;
        }
    }
}

// One inner product sum_i conj(v[i]) w[i] by one workgroup into out[0..1].  Thread t adds elements
// t, t + 256, ... in increasing order, eight elements' loads issued together.  Each term is
// re v.re w.re + v.im w.im = fma(v.re, w.re, v.im w.im), im fma(v.re, w.im, -(v.im w.re)), written
// out so that the rounding does not depend on which product the compiler fuses.
__device__ __forceinline__ void dot_block(const cplx* v, const cplx* ww, int n, double* sre, double* sim,
                                          double* out) {
    const int t = threadIdx.x;
    double re = 0, im = 0;
    for (int i0 = t; i0 < n; i0 += 8 * 256) {
        cplx a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 256 * u < n) {
                a[u] = v[i0 + 256 * u];
                b[u] = ww[i0 + 256 * u];
            }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 256 * u < n) {
                const double pr = a[u].im * b[u].im, pi = a[u].im * b[u].re;
                re += fma(a[u].re, b[u].re, pr);
                im += fma(a[u].re, b[u].im, -pi);
            }
    }
    tree_sum_256(re, im, sre, sim, out);
}

// ---------------------------------------------------------------------------------------
// FP64 MFMA 16x16x4.  Operand layout (one wave): A(i,k) at lane i + 16k, B(k,j) at lane
// j + 16k; result element (4q + lane/16, lane%16) in component q of the v4d accumulator.
// A complex product uses four real MFMAs into four independent accumulators.
// ---------------------------------------------------------------------------------------
struct cacc {
    v4d rr, ii, ri, ir;  // re*re, im*im, re*im, im*re
};

__device__ __forceinline__ void cacc_zero(cacc& a) {
    a.rr = v4d{0, 0, 0, 0}; a.ii = a.rr; a.ri = a.rr; a.ir = a.rr;
}

__device__ __forceinline__ void cmfma(cacc& acc, cplx a, cplx b) {
    acc.rr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, acc.rr, 0, 0, 0);
    acc.ii = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, acc.ii, 0, 0, 0);
    acc.ri = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, acc.ri, 0, 0, 0);
    acc.ir = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, acc.ir, 0, 0, 0);
}

// element q of the accumulated complex tile
__device__ __forceinline__ cplx cacc_get(const cacc& a, int q) {
    return cmake(a.rr[q] - a.ii[q], a.ri[q] + a.ir[q]);
}

// ---------------------------------------------------------------------------------------
// source s(t) = 2|t| J(t^2) from the piecewise Chebyshev table (Clenshaw recurrence)
// ---------------------------------------------------------------------------------------
#include "rcmps_tables.h"

__device__ __forceinline__ double rc_source(double t, const double* __restrict__ cheb) {
    double a = fabs(t);
    if (a >= RC_SOURCE_TMAX) return 0.0;
    int seg = (int)(a * (RC_SOURCE_SEGMENTS / RC_SOURCE_TMAX));
    if (seg > RC_SOURCE_SEGMENTS - 1) seg = RC_SOURCE_SEGMENTS - 1;
    double z = 2.0 * a * (RC_SOURCE_SEGMENTS / RC_SOURCE_TMAX) - 2.0 * seg - 1.0;
    const double* c = cheb + seg * (RC_SOURCE_DEGREE + 1);
    double b1 = 0.0, b2 = 0.0;
    for (int j = RC_SOURCE_DEGREE; j > 0; --j) {
        double b0 = 2.0 * z * b1 - b2 + c[j];
        b2 = b1;
        b1 = b0;
    }
    return z * b1 - b2 + c[0];
}
