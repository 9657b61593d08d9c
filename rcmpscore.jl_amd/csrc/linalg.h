// Host-side declarations of the dense complex kernels (linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

struct GemmArgs {
    int Dp, D, nbatch, opA, opB, alpha_stride;
    long long strideA, strideB, strideC, strideD;   // in complex elements
    const double* A;
    const double* B;
    double* C;
    const double* alpha;    // per-batch real weight (index batch*alpha_stride) or null
    double beta;
    double alpha_scale;     // scalar factor on the product; 0 means 1
    const double* addD;     // optional: C += dscale * dalpha_b * D_b
    const double* dalpha;
    double dscale;
    double add_identity;    // optional: C += add_identity * dalpha_b * I (on the D x D block)
};

__global__ void rc_zgemm_batched(GemmArgs g);
template <int WPB>
__global__ void rc_zgemm_batched_w(GemmArgs g);
__global__ void rc_zgemm_batched_lds(GemmArgs g);
__global__ void rc_zgemm_reduce_partial(GemmArgs g, double* partial, int chunk);
__global__ void rc_sum_partials(const double* partial, int nchunks, int n, double scale, double beta, double* C);
__global__ void rc_sum_partials_seg(const double* partial, int nchunks, int n, double scale, double beta, double* C,
                                    long long sC);
__global__ void rc_lincomb(const double* X, long long stride, const double* coef, int ncoef, int n, double beta,
                           double* out);
__global__ void rc_lincomb_n(const double* X, long long stride, long long sg, const double* coef, int ncoef, int n,
                             double beta, double* out, long long so);
__global__ void rc_axpby(int n, double ar, double ai, const double* X, double br, double bi, const double* Y,
                         double* out);
__global__ void rc_herm(int Dp, int mode, double scale, const double* X, double* out);
__global__ void rc_sylvester_diag(int Dp, int D, int adjoint, double shift, const double* lam, const double* Y,
                                  double* Z);
__global__ void rc_colmajor_in(int D, int Dp, const double* src, double* dst);
__global__ void rc_colmajor_out(int D, int Dp, const double* src, double* dst);
__global__ void rc_sylvester_tri(int Dp, int D, int adjoint, double s_re, double s_im, const double* T,
                                 const double* Y, double* Z);
size_t rc_sylvester_tri_lds(int D);   // dynamic LDS bytes of rc_sylvester_tri (one workgroup of 256)
#define RC_SYLVESTER_TRI_MAX_D 64
__global__ void rc_dots(const double* V, long long stride, int n, const double* w, double* out);
// Arnoldi steps of up to four bases in lockstep: basis b's step j orthogonalises w against V_0..V_j
// (columns h1 and h2 of the two Gram-Schmidt passes, h0 = |w|^2 before and hn after them) and
// writes V_{j+1}; vectors of n complex entries at stride n
struct RcArnBasis {
    double* V;
    double* w;
    const double* wl;   // w again, the left operand of |w|^2 (the products then compile as for any pair)
    double *h0, *h1, *h2, *hn;
    int j, pad;
};
struct RcArnBatch {
    int n, nb;
    RcArnBasis b[4];
};
__global__ void rc_arn_dots(RcArnBatch B, int pass);
__global__ void rc_arn_subtract(RcArnBatch B, int pass);
__global__ void rc_arn_normalize(RcArnBatch B);
__global__ void rc_subtract_combo(const double* V, long long stride, int nk, const double* h, int n, double* w);
__global__ void rc_add_combo(const double* V, long long stride, int nk, const double* y, int n, double* x);
__global__ void rc_trace_prod(int Dp, const double* A, const double* B, double* out);
__global__ void rc_add_identity(int Dp, int D, const double* cptr, double scale, double* X);
__global__ void rc_sub_scaled(int n, const double* cptr, double dre, double dim, const double* Z, double* X);
__global__ void rc_normalize(int n, const double* w, const double* nrm2, double* out);
__global__ void rc_neg_add_identity(int Dp, int D, const double* cptr, double scale, double* Y);
__global__ void rc_basis_combine(const double* V, long long stride, int k, const double* Y, int nnodes, int len,
                                 double* X);
__global__ void rc_shifted_residual(const double* X, const double* LX, const double* rhs, const double* z, int len,
                                    double* out);
