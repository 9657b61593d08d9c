// Host-side declarations of the excitation kernels (excitation_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

// coefficients of one DOP853 combination (up to 13 stage vectors)
struct RcStageCoef {
    double a[13];
    int nk;
};

// arguments of rc_quartic_rhs (the state and derivative hold 4 blocks of 5 na matrices: forward x,
// forward y, adjoint a0, adjoint a1; Vrep/Wrep hold v, w of tangent a at (5 a) Dp^2)
struct RcQuarticArgs {
    int Dp, na;
    const double* U;
    double* F;
    const double *Q, *Qh, *Qs, *Qhs, *R, *Rh, *Vrep, *Wrep;
    const double* coef;   // device [jac, jt] when non-null, else the values below
    double jac, jt;
};

__global__ void rc_quartic_rhs(RcQuarticArgs g);
__global__ void rc_sylvester_diag_c(int Dp, int D, int adjoint, double s_re, double s_im, const double* lam,
                                    const double* Y, double* Z);
__global__ void rc_mdots(const double* V, long long sb, long long sv, int nk, const double* W, long long sw, int n,
                         double* out, int ldo);
__global__ void rc_msubtract(const double* V, long long sb, long long sv, int nk, const double* h, int ldh, int n,
                             double* W, long long sw);
__global__ void rc_mnormalize(int n, const double* W, long long sw, const double* nrm2, int ldn, double* out,
                              long long so);
__global__ void rc_stage(long long N, const double* y, const double* K, long long sk, RcStageCoef c, double h,
                         double* out);
__global__ void rc_step_error(int len, const double* y, const double* ynew, const double* K, long long sk,
                              RcStageCoef e5, RcStageCoef e3, double atol, double rtol, double* out);
__global__ void rc_trace_prod_batched(int Dp, const double* A, long long sa, const double* B, long long sb,
                                      double* out);
__global__ void rc_axpy_c(int n, double c_re, double c_im, const double* Z, double* X);
