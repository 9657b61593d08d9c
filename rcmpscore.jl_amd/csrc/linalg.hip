// Dense complex FP64 building blocks on MFMA for Dp x Dp matrices (Dp a multiple of 16):
// strided-batched products with per-batch real weights, deterministic batch reductions
// (the pullback's sums over quadrature nodes), and the elementwise / Krylov helpers.
#include "device_common.h"
#include "linalg.h"

namespace {

// one 16x16 output tile per workgroup, 4 waves split the k-steps; operands read directly
// into MFMA fragments (op 0: as stored, op 1: conjugate transpose)
__device__ __forceinline__ cplx load_op(const cplx* M, int Dp, int op, int r, int c) {
    if (!op) return M[r * Dp + c];
    cplx z = M[c * Dp + r];
    return cmake(z.re, -z.im);
}

// A wave's share of one tile product: k-steps ks = wave, wave + 4, ... (at most four for Dp <= 64).
// All fragments are loaded before the first MFMA, so the wave waits for one round trip instead of
// one per k-step; the MFMAs then run in increasing ks, the order of the plain loop.
struct TileFrags {
    cplx a[4], b[4];
    double alpha;
};

__device__ __forceinline__ void tile_load(TileFrags& f, const cplx* A, int opA, const cplx* B, int opB, int Dp,
                                          int bi, int bj, int wave, int lane) {
    const int ks_total = Dp / 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ks = wave + 4 * j;
        if (ks < ks_total) {
            const int kk = 4 * ks + (lane >> 4);
            f.a[j] = load_op(A, Dp, opA, 16 * bi + (lane & 15), kk);
            f.b[j] = load_op(B, Dp, opB, kk, 16 * bj + (lane & 15));
        }
    }
}

__device__ __forceinline__ void tile_fma(cacc& acc, const TileFrags& f, int Dp, int wave) {
    const int ks_total = Dp / 4;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (wave + 4 * j < ks_total) cmfma(acc, f.a[j], f.b[j]);
}

__device__ __forceinline__ void tile_mma(cacc& acc, const cplx* A, int opA, const cplx* B, int opB,
                                         int Dp, int bi, int bj, int wave, int lane) {
    if (Dp <= 64) {
        TileFrags f;
        tile_load(f, A, opA, B, opB, Dp, bi, bj, wave, lane);
        tile_fma(acc, f, Dp, wave);
        return;
    }
    const int ks_total = Dp / 4;
    for (int ks = wave; ks < ks_total; ks += 4) {
        const int kk = 4 * ks + (lane >> 4);
        cplx a = load_op(A, Dp, opA, 16 * bi + (lane & 15), kk);
        cplx b = load_op(B, Dp, opB, kk, 16 * bj + (lane & 15));
        cmfma(acc, a, b);
    }
}

__device__ __forceinline__ cplx reduce4(const cacc& acc, cplx* red, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 256 + (4 * q + (lane >> 4)) * 16 + (lane & 15)] = cacc_get(acc, q);
    __syncthreads();
    const int t = threadIdx.x;
    cplx s = cadd(cadd(cadd(red[t], red[256 + t]), red[512 + t]), red[768 + t]);
    __syncthreads();
    return s;
}

// The (tile, batch) of a workgroup in a grid (tiles, batches).  Workgroups go to the 8 XCDs
// round-robin in linear order (x fastest); when the batches are a multiple of 8, the tiles of one
// batch are given to workgroups of one XCD, so its L2 serves the batch's operand panels to all of
// them instead of every XCD reading them from HBM.  A bijection: each (tile, batch) is computed
// exactly once, by the same arithmetic, so only the placement changes.
__device__ __forceinline__ void xcd_tile_batch(int& tile, int& batch) {
    const int nt = gridDim.x;
    if ((gridDim.y & 7) == 0) {
        const int lin = blockIdx.x + nt * blockIdx.y;
        const int slot = lin >> 3;
        batch = (lin & 7) + 8 * (slot / nt);
        tile = slot % nt;
    } else {
        tile = blockIdx.x;
        batch = blockIdx.y;
    }
}

}  // namespace

// C_b = alpha_b op(A_b) op(B_b) + beta C_b ; grid (NB*NB, nbatch)
__global__ void __launch_bounds__(256) rc_zgemm_batched(GemmArgs g) {
    __shared__ cplx red[1024];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int tile, batch;
    xcd_tile_batch(tile, batch);
    const int nb = g.Dp / 16, bi = tile / nb, bj = tile % nb;
    const cplx* A = reinterpret_cast<const cplx*>(g.A) + (long long)batch * g.strideA;
    const cplx* B = reinterpret_cast<const cplx*>(g.B) + (long long)batch * g.strideB;
    cplx* C = reinterpret_cast<cplx*>(g.C) + (long long)batch * g.strideC;
    // the epilogue's operands (this thread's element of C and of the added matrix, the weights)
    // are loaded with the fragments, so their round trip overlaps the product's
    const int r = 16 * bi + (tid >> 4), c = 16 * bj + (tid & 15);
    const double al = (g.alpha ? g.alpha[batch * g.alpha_stride] : 1.0) * (g.alpha_scale != 0.0 ? g.alpha_scale : 1.0);
    const double dl = g.dalpha ? g.dalpha[batch * g.alpha_stride] : 1.0;
    const cplx cin = g.beta != 0.0 ? C[r * g.Dp + c] : cmake(0, 0);
    const cplx din = g.addD ? (reinterpret_cast<const cplx*>(g.addD) + (long long)batch * g.strideD)[r * g.Dp + c]
                            : cmake(0, 0);
    cacc acc;
    cacc_zero(acc);
    tile_mma(acc, A, g.opA, B, g.opB, g.Dp, bi, bj, wave, lane);
    cplx v = reduce4(acc, red, wave, lane);
    cplx out = cscale(al, v);
    if (g.beta != 0.0) out = cadd(out, cscale(g.beta, cin));
    if (g.addD) out = cadd(out, cscale(dl * g.dscale, din));
    if (g.add_identity != 0.0 && r == c && r < g.D) out.re += dl * g.add_identity;
    C[r * g.Dp + c] = out;
}

// The same product with one 16x16 output tile per wave and WPB waves per workgroup (grid
// (ceil(NB*NB / WPB), nbatch)): a wave runs all k-steps itself, keeping the four partial tiles of
// the k-steps ks = w (mod 4) that rc_zgemm_batched's waves w = 0..3 accumulate, and sums them in
// the same order, so every variant produces the same bits.  No LDS and no barrier.
template <int WPB>
__global__ void __launch_bounds__(64 * WPB) rc_zgemm_batched_w(GemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb = g.Dp / 16, tile = blockIdx.x * WPB + wave, batch = blockIdx.y;
    if (tile >= nb * nb) return;
    const int bi = tile / nb, bj = tile % nb;
    const cplx* A = reinterpret_cast<const cplx*>(g.A) + (long long)batch * g.strideA;
    const cplx* B = reinterpret_cast<const cplx*>(g.B) + (long long)batch * g.strideB;
    cplx* C = reinterpret_cast<cplx*>(g.C) + (long long)batch * g.strideC;
    cacc acc[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) cacc_zero(acc[w]);
    const int ks_total = g.Dp / 4;   // a multiple of 4
    for (int ks = 0; ks < ks_total; ks += 4) {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int kk = 4 * (ks + w) + (lane >> 4);
            cmfma(acc[w], load_op(A, g.Dp, g.opA, 16 * bi + (lane & 15), kk),
                  load_op(B, g.Dp, g.opB, kk, 16 * bj + (lane & 15)));
        }
    }
    const double al = (g.alpha ? g.alpha[batch * g.alpha_stride] : 1.0) * (g.alpha_scale != 0.0 ? g.alpha_scale : 1.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const cplx v = cadd(cadd(cadd(cacc_get(acc[0], q), cacc_get(acc[1], q)), cacc_get(acc[2], q)),
                            cacc_get(acc[3], q));
        const int r = 16 * bi + 4 * q + (lane >> 4), c = 16 * bj + (lane & 15);
        cplx out = cscale(al, v);
        if (g.beta != 0.0) out = cadd(out, cscale(g.beta, C[r * g.Dp + c]));
        if (g.addD) {
            const cplx* Dm = reinterpret_cast<const cplx*>(g.addD) + (long long)batch * g.strideD;
            const double dl = g.dalpha ? g.dalpha[batch * g.alpha_stride] : 1.0;
            out = cadd(out, cscale(dl * g.dscale, Dm[r * g.Dp + c]));
        }
        if (g.add_identity != 0.0 && r == c && r < g.D) {
            const double il = g.dalpha ? g.dalpha[batch * g.alpha_stride] : 1.0;
            out.re += il * g.add_identity;
        }
        C[r * g.Dp + c] = out;
    }
}
template __global__ void rc_zgemm_batched_w<1>(GemmArgs g);
template __global__ void rc_zgemm_batched_w<4>(GemmArgs g);

// A 32x32 output block per workgroup of four waves, one 16x16 quadrant per wave (grid
// (ceil(NB/2)^2, nbatch), dynamic LDS 2 * 32 * (Dp + 1) complex): the block's 32-row panel of op(A)
// and 32-column panel of op(B) are staged in LDS once with coalesced loads and shared by the
// waves (each operand element is read once per block instead of once per 16x16 tile).  The k-step
// classes and their order are those of rc_zgemm_batched, so the bits are the same.
__global__ void __launch_bounds__(256) rc_zgemm_batched_lds(GemmArgs g) {
    extern __shared__ cplx panel[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int Dp = g.Dp, nb = Dp / 16, nb2 = (nb + 1) / 2, ld = Dp + 1;
    const int bI = blockIdx.x / nb2, bJ = blockIdx.x % nb2, batch = blockIdx.y;
    const int rows = min(32, Dp - 32 * bI), cols = min(32, Dp - 32 * bJ);   // 16 or 32
    const cplx* A = reinterpret_cast<const cplx*>(g.A) + (long long)batch * g.strideA;
    const cplx* B = reinterpret_cast<const cplx*>(g.B) + (long long)batch * g.strideB;
    cplx* C = reinterpret_cast<cplx*>(g.C) + (long long)batch * g.strideC;
    cplx* Ap = panel;             // Ap[r * ld + k] = op(A)(32 bI + r, k)
    cplx* Bp = panel + 32 * ld;   // Bp[c * ld + k] = op(B)(k, 32 bJ + c)
    // op 0 reads A row-major along k; op 1 (conjugate transpose) reads A along the panel's rows
    for (int i = tid; i < rows * Dp; i += 256) {
        if (!g.opA) {
            const int r = i / Dp, k = i % Dp;
            Ap[r * ld + k] = A[(32 * bI + r) * Dp + k];
        } else {
            const int k = i / rows, r = i % rows;
            const cplx z = A[k * Dp + 32 * bI + r];
            Ap[r * ld + k] = cmake(z.re, -z.im);
        }
    }
    for (int i = tid; i < cols * Dp; i += 256) {
        if (!g.opB) {
            const int k = i / cols, c = i % cols;
            Bp[c * ld + k] = B[k * Dp + 32 * bJ + c];
        } else {
            const int c = i / Dp, k = i % Dp;
            const cplx z = B[(32 * bJ + c) * Dp + k];
            Bp[c * ld + k] = cmake(z.re, -z.im);
        }
    }
    __syncthreads();
    const int qi = wave >> 1, qj = wave & 1;
    if (16 * qi >= rows || 16 * qj >= cols) return;
    cacc acc[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) cacc_zero(acc[w]);
    const int ks_total = Dp / 4;
    const cplx* Ar = Ap + (16 * qi + (lane & 15)) * ld;
    const cplx* Bc = Bp + (16 * qj + (lane & 15)) * ld;
    for (int ks = 0; ks < ks_total; ks += 4) {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int kk = 4 * (ks + w) + (lane >> 4);
            cmfma(acc[w], Ar[kk], Bc[kk]);
        }
    }
    const double al = (g.alpha ? g.alpha[batch * g.alpha_stride] : 1.0) * (g.alpha_scale != 0.0 ? g.alpha_scale : 1.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const cplx v = cadd(cadd(cadd(cacc_get(acc[0], q), cacc_get(acc[1], q)), cacc_get(acc[2], q)),
                            cacc_get(acc[3], q));
        const int r = 32 * bI + 16 * qi + 4 * q + (lane >> 4), c = 32 * bJ + 16 * qj + (lane & 15);
        cplx out = cscale(al, v);
        if (g.beta != 0.0) out = cadd(out, cscale(g.beta, C[r * Dp + c]));
        if (g.addD) {
            const cplx* Dm = reinterpret_cast<const cplx*>(g.addD) + (long long)batch * g.strideD;
            const double dl = g.dalpha ? g.dalpha[batch * g.alpha_stride] : 1.0;
            out = cadd(out, cscale(dl * g.dscale, Dm[r * Dp + c]));
        }
        if (g.add_identity != 0.0 && r == c && r < g.D) {
            const double il = g.dalpha ? g.dalpha[batch * g.alpha_stride] : 1.0;
            out.re += il * g.add_identity;
        }
        C[r * Dp + c] = out;
    }
}

// partial[chunk] = sum_{b in chunk} alpha_b op(A_b) op(B_b) ; grid (NB*NB, nchunks)
// The next batch's fragments and weight are loaded while the current batch's MFMAs run; the
// batches are still accumulated one after another in increasing b.
__global__ void __launch_bounds__(256) rc_zgemm_reduce_partial(GemmArgs g, double* partial, int chunk) {
    __shared__ cplx red[1024];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int tile, ci;
    xcd_tile_batch(tile, ci);
    const int nb = g.Dp / 16, bi = tile / nb, bj = tile % nb;
    const int b0 = ci * chunk, b1 = min(g.nbatch, b0 + chunk);
    const double ascale = g.alpha_scale != 0.0 ? g.alpha_scale : 1.0;
    auto fetch = [&](TileFrags& f, int batch) {
        const cplx* A = reinterpret_cast<const cplx*>(g.A) + (long long)batch * g.strideA;
        const cplx* B = reinterpret_cast<const cplx*>(g.B) + (long long)batch * g.strideB;
        f.alpha = g.alpha ? g.alpha[batch * g.alpha_stride] : 1.0;
        tile_load(f, A, g.opA, B, g.opB, g.Dp, bi, bj, wave, lane);
    };
    cacc acc;
    cacc_zero(acc);
    if (g.Dp <= 64 && b0 < b1) {
        TileFrags cur, nxt;
        fetch(cur, b0);
        for (int batch = b0; batch < b1; ++batch) {
            if (batch + 1 < b1) fetch(nxt, batch + 1);
            const double al = cur.alpha * ascale;
            cacc part;
            cacc_zero(part);
            tile_fma(part, cur, g.Dp, wave);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc.rr[q] += al * part.rr[q];
                acc.ii[q] += al * part.ii[q];
                acc.ri[q] += al * part.ri[q];
                acc.ir[q] += al * part.ir[q];
            }
            cur = nxt;
        }
    } else {
        for (int batch = b0; batch < b1; ++batch) {
            const cplx* A = reinterpret_cast<const cplx*>(g.A) + (long long)batch * g.strideA;
            const cplx* B = reinterpret_cast<const cplx*>(g.B) + (long long)batch * g.strideB;
            const double al = (g.alpha ? g.alpha[batch * g.alpha_stride] : 1.0) * ascale;
            cacc part;
            cacc_zero(part);
            tile_mma(part, A, g.opA, B, g.opB, g.Dp, bi, bj, wave, lane);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc.rr[q] += al * part.rr[q];
                acc.ii[q] += al * part.ii[q];
                acc.ri[q] += al * part.ri[q];
                acc.ir[q] += al * part.ir[q];
            }
        }
    }
    cplx v = reduce4(acc, red, wave, lane);
    const int r = 16 * bi + (tid >> 4), c = 16 * bj + (tid & 15);
    reinterpret_cast<cplx*>(partial)[(size_t)ci * g.Dp * g.Dp + r * g.Dp + c] = v;
}

// C = beta C + scale * sum_chunks partial (fixed order)
__device__ __forceinline__ void sum_partials_elem(const double* partial, int nchunks, int n, double scale,
                                                  double beta, double* C, int i) {
    const cplx* p = reinterpret_cast<const cplx*>(partial);
    cplx s = cmake(0, 0);
    // eight chunks' loads in flight, added in increasing k
    for (int k0 = 0; k0 < nchunks; k0 += 8) {
        cplx x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nchunks) x[u] = p[(size_t)(k0 + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nchunks) s = cadd(s, x[u]);
    }
    cplx* Cc = reinterpret_cast<cplx*>(C);
    cplx out = cscale(scale, s);
    if (beta != 0.0) out = cadd(out, cscale(beta, Cc[i]));
    Cc[i] = out;
}

__global__ void rc_sum_partials(const double* partial, int nchunks, int n, double scale, double beta, double* C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sum_partials_elem(partial, nchunks, n, scale, beta, C, i);
}

// rc_sum_partials per segment (blockIdx.y = s): the nchunks partials of segment s into C + s sC
__global__ void rc_sum_partials_seg(const double* partial, int nchunks, int n, double scale, double beta, double* C,
                                    long long sC) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long sg = blockIdx.y;
    sum_partials_elem(partial + 2 * sg * nchunks * n, nchunks, n, scale, beta, C + 2 * sg * sC, i);
}

// out = sum_k coef[k] * X_k   (X_k = X + k*stride), n complex elements; coef real
// Sixteen terms' loads are issued together before they are added in order k = 0, 1, ..., so the
// node sums of the pullback overlap their memory latency and keep the sequential sum's bits.
__device__ __forceinline__ void lincomb_elem(const double* X, long long stride, const double* coef, int ncoef, double beta,
                                             double* out, int i) {
    const cplx* x = reinterpret_cast<const cplx*>(X);
    cplx s = cmake(0, 0);
    int k = 0;
    for (; k + 16 <= ncoef; k += 16) {
        cplx v[16];
        double c[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            v[q] = x[(long long)(k + q) * stride + i];
            c[q] = coef[k + q];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) s = cadd(s, cscale(c[q], v[q]));
    }
    for (; k < ncoef; ++k) s = cadd(s, cscale(coef[k], x[(long long)k * stride + i]));
    cplx* o = reinterpret_cast<cplx*>(out);
    if (beta != 0.0) s = cadd(s, cscale(beta, o[i]));
    o[i] = s;
}

__global__ void rc_lincomb(const double* X, long long stride, const double* coef, int ncoef, int n,
                           double beta, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lincomb_elem(X, stride, coef, ncoef, beta, out, i);
}

// rc_lincomb for several groups at once (blockIdx.y = g): X + g sg, out + g so (complex strides)
__global__ void rc_lincomb_n(const double* X, long long stride, long long sg, const double* coef, int ncoef, int n,
                             double beta, double* out, long long so) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long g = blockIdx.y;
    lincomb_elem(X + 2 * g * sg, stride, coef, ncoef, beta, out + 2 * g * so, i);
}

// elementwise: out = a*X + b*Y (complex scalars)
__global__ void rc_axpby(int n, double ar, double ai, const double* X, double br, double bi, const double* Y,
                         double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cplx* x = reinterpret_cast<const cplx*>(X);
    const cplx* y = reinterpret_cast<const cplx*>(Y);
    cplx v = cmul(cmake(ar, ai), x[i]);
    if (Y) v = cadd(v, cmul(cmake(br, bi), y[i]));
    reinterpret_cast<cplx*>(out)[i] = v;
}

// out = (X + X^+)/2 scaled; or out = X^+ (mode 1); Dp x Dp
__global__ void rc_herm(int Dp, int mode, double scale, const double* X, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Dp * Dp) return;
    const int r = i / Dp, c = i % Dp;
    const cplx* x = reinterpret_cast<const cplx*>(X);
    cplx a = x[r * Dp + c], bt = cconj(x[c * Dp + r]);
    cplx v = mode == 0 ? cscale(0.5 * scale, cadd(a, bt)) : cscale(scale, bt);
    reinterpret_cast<cplx*>(out)[i] = v;
}

// Boundary copies of a column-major D x D matrix (element (i, j) at i + j D, Julia's layout):
// into the padded row-major Dp x Dp working layout (zero padding), and back out.
__global__ void rc_colmajor_in(int D, int Dp, const double* src, double* dst) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Dp * Dp) return;
    const int i = e / Dp, j = e % Dp;
    reinterpret_cast<cplx*>(dst)[e] =
        (i < D && j < D) ? reinterpret_cast<const cplx*>(src)[(size_t)j * D + i] : cmake(0, 0);
}

__global__ void rc_colmajor_out(int D, int Dp, const double* src, double* dst) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D * D) return;
    const int i = e % D, j = e / D;   // dst element e = (i, j) column-major
    reinterpret_cast<cplx*>(dst)[e] = reinterpret_cast<const cplx*>(src)[(size_t)i * Dp + j];
}

// Z_ij = -Y_ij / den_ij   (den = lam_i + conj(lam_j) - shift, or conj(lam_i) + lam_j - shift)
__global__ void rc_sylvester_diag(int Dp, int D, int adjoint, double shift, const double* lam, const double* Y,
                                  double* Z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Dp * Dp) return;
    const int r = i / Dp, c = i % Dp;
    cplx* z = reinterpret_cast<cplx*>(Z);
    if (r >= D || c >= D) { z[i] = cmake(0, 0); return; }
    const cplx* l = reinterpret_cast<const cplx*>(lam);
    cplx den = adjoint ? cadd(cconj(l[r]), l[c]) : cadd(l[r], cconj(l[c]));
    den.re -= shift;
    cplx y = reinterpret_cast<const cplx*>(Y)[i];
    double m = cabs2(den);
    cplx q = cmake((y.re * den.re + y.im * den.im) / m, (y.im * den.re - y.re * den.im) / m);
    z[i] = cmake(-q.re, -q.im);
}

// The same Sylvester solve in the Schur basis of Q (Q = Z T Z^+, T upper triangular), for a Q whose
// eigenvectors are ill-conditioned (defective or nearly so):
//   forward  T X + X T^+ - s X = -Y,   adjoint  T^+ X + X T - s X = -Y,   s = s_re + i s_im,
// by the Bartels-Stewart recurrence (src/transfer.jl:50-68 solves the same triangular system with
// LAPACK's trsyl).  Forward, X_ij needs X_kj for k > i and X_ik for k > j; adjoint, k < i and k < j:
// the elements of one anti-diagonal i + j = const are independent, so the sweep runs over the 2D - 1
// anti-diagonals (from the bottom-right corner forward, from the top-left corner for the adjoint).
// One workgroup of 256 threads, T and X in LDS (2 D^2 complex, D <= 64); four lanes per element
// split both inner sums (k = part, part + 4, ...) and add them by shuffles.
__global__ void __launch_bounds__(256) rc_sylvester_tri(int Dp, int D, int adjoint, double s_re, double s_im,
                                                        const double* T, const double* Y, double* Z) {
    extern __shared__ cplx syl_lds[];
    cplx* sT = syl_lds;
    cplx* sX = syl_lds + D * D;
    const int tid = threadIdx.x;
    const cplx* t = reinterpret_cast<const cplx*>(T);
    const cplx* y = reinterpret_cast<const cplx*>(Y);
    for (int e = tid; e < D * D; e += 256) {
        sT[e] = t[(e / D) * Dp + e % D];
        sX[e] = cmake(0, 0);
    }
    __syncthreads();
    const int nd = 2 * D - 1, part = tid & 3;
    for (int step = 0; step < nd; ++step) {
        const int sd = adjoint ? step : nd - 1 - step;
        const int i_lo = sd > D - 1 ? sd - (D - 1) : 0, i_hi = sd < D - 1 ? sd : D - 1;
        for (int e0 = 0; e0 <= i_hi - i_lo; e0 += 64) {
            const int e = e0 + (tid >> 2);
            const bool live = e <= i_hi - i_lo;
            const int i = i_lo + e, j = sd - i;
            double re = 0.0, im = 0.0;
            if (live) {
                if (!adjoint) {
                    for (int k = i + 1 + part; k < D; k += 4) {     // T_ik X_kj
                        const cplx a = sT[i * D + k], b = sX[k * D + j];
                        re += a.re * b.re - a.im * b.im;
                        im += a.re * b.im + a.im * b.re;
                    }
                    for (int k = j + 1 + part; k < D; k += 4) {     // X_ik conj(T_jk)
                        const cplx a = sX[i * D + k], b = sT[j * D + k];
                        re += a.re * b.re + a.im * b.im;
                        im += a.im * b.re - a.re * b.im;
                    }
                } else {
                    for (int k = part; k < i; k += 4) {             // conj(T_ki) X_kj
                        const cplx a = sT[k * D + i], b = sX[k * D + j];
                        re += a.re * b.re + a.im * b.im;
                        im += a.re * b.im - a.im * b.re;
                    }
                    for (int k = part; k < j; k += 4) {             // X_ik T_kj
                        const cplx a = sX[i * D + k], b = sT[k * D + j];
                        re += a.re * b.re - a.im * b.im;
                        im += a.re * b.im + a.im * b.re;
                    }
                }
            }
            re += __shfl_xor(re, 1);
            im += __shfl_xor(im, 1);
            re += __shfl_xor(re, 2);
            im += __shfl_xor(im, 2);
            if (live && part == 0) {
                const cplx ti = sT[i * D + i], tj = sT[j * D + j];
                cplx den = adjoint ? cadd(cconj(ti), tj) : cadd(ti, cconj(tj));
                den.re -= s_re;
                den.im -= s_im;
                const cplx yy = y[i * Dp + j];
                const double nr = yy.re + re, ni = yy.im + im, m = cabs2(den);
                sX[i * D + j] = cmake(-(nr * den.re + ni * den.im) / m, -(ni * den.re - nr * den.im) / m);
            }
        }
        __syncthreads();
    }
    cplx* z = reinterpret_cast<cplx*>(Z);
    for (int e = tid; e < Dp * Dp; e += 256) {
        const int r = e / Dp, c = e % Dp;
        z[e] = (r < D && c < D) ? sX[r * D + c] : cmake(0, 0);
    }
}

size_t rc_sylvester_tri_lds(int D) { return (size_t)2 * D * D * sizeof(cplx); }

// Block-parallel inner products: out[k] = sum_i conj(V_k[i]) w[i], one workgroup per k.
__global__ void __launch_bounds__(256) rc_dots(const double* V, long long stride, int n, const double* w,
                                               double* out) {
    __shared__ double sre[256], sim[256];
    const int k = blockIdx.x;
    dot_block(reinterpret_cast<const cplx*>(V) + (size_t)k * stride, reinterpret_cast<const cplx*>(w), n, sre, sim,
              out + 2 * k);
}

// element i of w -= sum_k h_k V_k
__device__ __forceinline__ void subtract_elem(const double* V, long long stride, int nk, const double* h, int i,
                                              double* w) {
    const cplx* v = reinterpret_cast<const cplx*>(V);
    const cplx* hh = reinterpret_cast<const cplx*>(h);
    cplx s = reinterpret_cast<cplx*>(w)[i];
    // eight vectors' loads in flight, subtracted in increasing k
    for (int k0 = 0; k0 < nk; k0 += 8) {
        cplx x[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nk) {
                x[u] = v[(size_t)(k0 + u) * stride + i];
                c[u] = hh[k0 + u];
            }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nk) s = csub(s, cmul(c[u], x[u]));
    }
    reinterpret_cast<cplx*>(w)[i] = s;
}

// w -= sum_k h_k V_k   (h complex, host-provided array in device memory)
__global__ void rc_subtract_combo(const double* V, long long stride, int nk, const double* h, int n, double* w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    subtract_elem(V, stride, nk, h, i, w);
}

// ---- Arnoldi steps of several bases in lockstep (blockIdx.y = basis); the arithmetic of each
// basis is that of rc_dots, rc_subtract_combo and rc_normalize, so each gives the same bits as
// its own step-by-step sequence.
// pass 1: blocks 0..j: h1[k] = <V_k, w>, block j + 1: h0 = <w, w>; pass 2: blocks 0..j: h2[k] = <V_k, w>
__global__ void __launch_bounds__(256) rc_arn_dots(RcArnBatch B, int pass) {
    __shared__ double sre[256], sim[256];
    const RcArnBasis& b = B.b[blockIdx.y];
    const int k = blockIdx.x;
    if (k > b.j + (pass == 1 ? 1 : 0)) return;
    const cplx* w = reinterpret_cast<const cplx*>(b.w);
    if (k == b.j + 1) {
        dot_block(reinterpret_cast<const cplx*>(b.wl), w, B.n, sre, sim, b.h0);
        return;
    }
    dot_block(reinterpret_cast<const cplx*>(b.V) + (size_t)k * B.n, w, B.n, sre, sim, (pass == 1 ? b.h1 : b.h2) + 2 * k);
}

// w -= sum_{k <= j} h_k V_k with the pass's projections
__global__ void rc_arn_subtract(RcArnBatch B, int pass) {
    const RcArnBasis& b = B.b[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    subtract_elem(b.V, B.n, b.j + 1, pass == 1 ? b.h1 : b.h2, i, b.w);
}

// hn = <w, w>, then V_{j+1} = w / sqrt(Re hn) (zero stays zero); one workgroup per basis
__global__ void __launch_bounds__(256) rc_arn_normalize(RcArnBatch B) {
    __shared__ double sre[256], sim[256];
    __shared__ double s_n;
    const RcArnBasis& b = B.b[blockIdx.y];
    const cplx* w = reinterpret_cast<const cplx*>(b.w);
    dot_block(reinterpret_cast<const cplx*>(b.wl), w, B.n, sre, sim, b.hn);
    if (threadIdx.x == 0) s_n = b.hn[0];   // this thread's own store, read back
    __syncthreads();
    const double nrm2 = s_n;
    const double sc = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 0.0;
    cplx* out = reinterpret_cast<cplx*>(b.V) + (size_t)(b.j + 1) * B.n;
    for (int i = threadIdx.x; i < B.n; i += 256) out[i] = cscale(sc, w[i]);
}

// x += sum_k y_k V_k
__global__ void rc_add_combo(const double* V, long long stride, int nk, const double* y, int n, double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cplx* v = reinterpret_cast<const cplx*>(V);
    const cplx* yy = reinterpret_cast<const cplx*>(y);
    cplx s = reinterpret_cast<cplx*>(x)[i];
    for (int k = 0; k < nk; ++k) s = cadd(s, cmul(yy[k], v[(size_t)k * stride + i]));
    reinterpret_cast<cplx*>(x)[i] = s;
}

// out[0] = trace(A B) for Dp x Dp (single workgroup).  Thread t adds the products of entries
// idx = t, t + 256, ... in increasing order, eight entries' loads issued together.
__global__ void __launch_bounds__(256) rc_trace_prod(int Dp, const double* A, const double* B, double* out) {
    __shared__ double sre[256], sim[256];
    const int t = threadIdx.x, n = Dp * Dp;
    const cplx* a = reinterpret_cast<const cplx*>(A);
    const cplx* b = reinterpret_cast<const cplx*>(B);
    double re = 0, im = 0;
    for (int i0 = t; i0 < n; i0 += 8 * 256) {
        cplx x[8], y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = i0 + 256 * u;
            if (idx < n) {
                const int i = idx / Dp, k = idx % Dp;
                x[u] = a[i * Dp + k];
                y[u] = b[k * Dp + i];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 256 * u < n) {
                cplx p = cmul(x[u], y[u]);
                re += p.re;
                im += p.im;
            }
    }
    tree_sum_256(re, im, sre, sim, out);
}

// X += (c_re + i c_im) I on the D x D block, c read from device memory at cptr (scaled)
__global__ void rc_add_identity(int Dp, int D, const double* cptr, double scale, double* X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    cplx* x = reinterpret_cast<cplx*>(X);
    x[i * Dp + i].re += scale * cptr[0];
    x[i * Dp + i].im += scale * cptr[1];
}

// X -= (t/den) * Z with t = cptr[0..1] (complex), den complex (host)
__global__ void rc_sub_scaled(int n, const double* cptr, double dre, double dim, const double* Z, double* X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx tt = cmake(cptr[0], cptr[1]);
    double m = dre * dre + dim * dim;
    cplx f = cmake((tt.re * dre + tt.im * dim) / m, (tt.im * dre - tt.re * dim) / m);
    cplx* x = reinterpret_cast<cplx*>(X);
    x[i] = csub(x[i], cmul(f, reinterpret_cast<const cplx*>(Z)[i]));
}

// out = w / sqrt(Re nrm2[0]) (nrm2 is a device <w, w>); zero stays zero
__global__ void rc_normalize(int n, const double* w, const double* nrm2, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = nrm2[0] > 0.0 ? 1.0 / sqrt(nrm2[0]) : 0.0;
    reinterpret_cast<cplx*>(out)[i] = cscale(s, reinterpret_cast<const cplx*>(w)[i]);
}

// y = -y + (scale * c) I on the D x D block, c complex in device memory
__global__ void rc_neg_add_identity(int Dp, int D, const double* cptr, double scale, double* Y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Dp * Dp) return;
    cplx* y = reinterpret_cast<cplx*>(Y);
    cplx v = cscale(-1.0, y[i]);
    const int r = i / Dp, c = i % Dp;
    if (r == c && r < D) {
        v.re += scale * cptr[0];
        v.im += scale * cptr[1];
    }
    y[i] = v;
}

// Krylov reconstruction X_i = sum_j Y[j + k i] V_j for nodes i (Y complex, one column per node);
// each workgroup holds 256 elements of 8 nodes so that V is read once per 8 nodes.
__global__ void __launch_bounds__(256) rc_basis_combine(const double* V, long long stride, int k, const double* Y,
                                                        int nnodes, int len, double* X) {
    __shared__ cplx ys[8][64];
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int i0 = blockIdx.y * 8;
    const cplx* v = reinterpret_cast<const cplx*>(V);
    const cplx* y = reinterpret_cast<const cplx*>(Y);
    cplx acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = cmake(0, 0);
    for (int j0 = 0; j0 < k; j0 += 64) {
        for (int t = threadIdx.x; t < 8 * 64; t += 256) {
            const int q = t / 64, jj = t % 64;
            ys[q][jj] = (i0 + q < nnodes && j0 + jj < k) ? y[(size_t)(j0 + jj) + (size_t)k * (i0 + q)] : cmake(0, 0);
        }
        __syncthreads();
        if (e < len) {
            const int jn = k - j0 < 64 ? k - j0 : 64;
            for (int jj = 0; jj < jn; ++jj) {
                const cplx vv = v[(size_t)(j0 + jj) * stride + e];
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] = cadd(acc[q], cmul(ys[q][jj], vv));
            }
        }
        __syncthreads();
    }
    if (e < len) {
        cplx* x = reinterpret_cast<cplx*>(X);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (i0 + q < nnodes) x[(size_t)(i0 + q) * len + e] = acc[q];
    }
}

// out[i] = |X_i - z_i LX_i - rhs|^2 over len elements, one workgroup per node i
__global__ void __launch_bounds__(256) rc_shifted_residual(const double* X, const double* LX, const double* rhs,
                                                           const double* z, int len, double* out) {
    __shared__ double s[256];
    const int i = blockIdx.x, t = threadIdx.x;
    const cplx* x = reinterpret_cast<const cplx*>(X) + (size_t)i * len;
    const cplx* lx = reinterpret_cast<const cplx*>(LX) + (size_t)i * len;
    const cplx* b = reinterpret_cast<const cplx*>(rhs);
    const double zi = z[i];
    double acc = 0;
    for (int e = t; e < len; e += 256) {
        const cplx r = csub(csub(x[e], cscale(zi, lx[e])), b[e]);
        acc += cabs2(r);
    }
    s[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) out[i] = s[0];
}
