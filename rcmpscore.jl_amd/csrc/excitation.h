// Host-side declarations of the excitation kernels (excitation_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// coefficients of one DOP853 combination (up to 13 stage vectors)
struct RcStageCoef {
    double a[13];
    int nk;
};

// Layout of the :phi^4: action ODE state of na tangents (five channels each, RC_QCH): the forward x
// channels, every tangent's forward y channels, every tangent's adjoint a0 channels, the adjoint a1
// channels.  x and a1 have no tangent-dependent operand and the same start (rho, I) for every
// tangent, so one copy serves all; rc_qmat gives the matrix index of block b, tangent a, channel ch
// (nb = RC_QCH na), rc_qmats the count.  The forward half (x, y) comes first.
#define RC_QCH 5
__host__ __device__ inline int rc_qmat(int b, int a, int ch, int nb) {
    return b == 0 ? ch : b == 1 ? RC_QCH + RC_QCH * a + ch : b == 2 ? RC_QCH + nb + RC_QCH * a + ch : RC_QCH + 2 * nb + ch;
}
__host__ __device__ inline int rc_qmats(int nb) { return 2 * RC_QCH + 2 * nb; }

// arguments of rc_quartic_rhs (the state and derivative in the layout above, rc_qmats(5 na) matrices;
// Vrep/Wrep hold v, w of tangent a at (5 a) Dp^2)
// The epilogue chains the DOP853 stages: epi = 1 writes the next stage input
// Unext = y + h (sum_{j<s} comb_j K_j + comb_s F) for this tile's elements (or the step end with
// the B weights), epi = 2 writes this tile's error sums sum |E5.K/sc|^2, sum |E3.K/sc|^2 with
// sc = atol + rtol max(|y|, |U|) (U is then the step end) into partial[2 (mi ntiles + tile)].
struct RcQuarticArgs {
    int Dp, na;
    const double* U;
    double* F;
    const double *Q, *Qh, *Qs, *Qhs, *R, *Rh, *Vrep, *Wrep;
    long long qstride;    // complex stride between the tangents' Qs (and Qhs); 0 when shared
    const double* coef;   // device [jac, jt] when non-null, else the values below
    double jac, jt;
    int epi, stage;
    const double* y;      // state at the step start
    const double* K;      // stage derivatives, K_j at K + j sk (complex stride)
    long long sk;
    RcStageCoef comb, comb3;
    double* Unext;
    const double* hp;     // device step size when non-null, else h
    double h, atol, rtol;
    double* partial;
    // Z2 structure: for R odd and K, rho even with parity blocks of zt whole tiles (0: none), and
    // every tangent of the launch of parity wpar (0 even, 1 odd), each channel has a definite parity
    // and the products run only the k-tiles of the one parity block that can be nonzero
    int zt, wpar;
    const int* active;    // device flag: the launch does nothing when it reads 0 (null: always runs)
};

__global__ void rc_quartic_rhs(RcQuarticArgs g);

// arguments of rc_pair_rhs: pair p's state at U + p ustride (matrix (channel c, block k) at (4 c + k)
// Dp^2), its derivative at F + p fstride; per-pair data (Q - ip, Q + ip, v = -R^+ w, w, vbra^+, u^+) at
// + p dstride and the overlap n at nov[2 p nstride]; epilogue fields as RcQuarticArgs
struct RcPairArgs {
    int Dp, npair, nch;
    const double* U;
    long long ustride;
    double* F;
    long long fstride;
    const double *Q, *Qh, *R, *Rh;
    const double *Qs, *Qb, *V, *W, *Vbh, *Uh;
    long long dstride;
    const double* nov;
    long long nstride;
    const double* coef;
    double jac, jt;
    int epi, stage;
    const double* y;
    const double* K;
    long long sk;
    RcStageCoef comb, comb3;
    double* Unext;
    const double* hp;
    double h, atol, rtol;
    double* partial;
    const int* active;    // as RcQuarticArgs::active
};

__global__ void rc_pair_rhs(RcPairArgs g);
__global__ void rc_res_scale(long long N, int nz, int per_basis, int adjoint, const double* Y, const double* mu,
                             const double* z, const double* s, double* out);
__global__ void rc_bcast_rhs(long long N, int nz, const double* src, long long ss, double* out);
__global__ void rc_poisson_scale(long long N, int Dp, int D, int adjoint, double s, int k0, const double* mu,
                                 const double* Y, double* out);
// up to four operands of rc_lindblad_multi, each with its direction
struct RcLindbladMulti {
    const double* X[4];
    double* out[4];
    int adj[4];
};
__global__ void rc_lindblad_multi(int Dp, const double* Q, const double* Qh, const double* R, const double* Rh,
                                  RcLindbladMulti M);
__global__ void rc_lindblad(int Dp, int adjoint, const double* Q, const double* Qh, const double* R, const double* Rh,
                            const double* X, double* out);
__global__ void rc_two_sided(int Dp, int D, const double* A, const double* X, const double* B, double* out, int sylvester,
                             int adjoint, const double* lam, double s_re, double s_im);
__global__ void rc_sylvester_diag_c(int Dp, int D, int adjoint, double s_re, double s_im, const double* lam,
                                    const double* Y, double* Z);
__global__ void rc_mdots(const double* V, long long sb, long long sv, int nk, const double* W, long long sw, int n,
                         double* out, int ldo);
__global__ void rc_msubtract(const double* V, long long sb, long long sv, int nk, const double* h, int ldh, int n,
                             double* W, long long sw);
__global__ void rc_mnormalize(int n, const double* W, long long sw, const double* nrm2, int ldn, double* out,
                              long long so);
__global__ void rc_stage(long long N, const double* y, const double* K, long long sk, RcStageCoef c, double h,
                         const double* hp, double* out, const int* active);

// Device-resident DOP853 step controller of the action ODE.  One attempt is rc_ode_stage (which
// first commits the previous attempt when it was accepted: y <- y_new, K_0 <- K_12, the sample of a
// reached quadrature node; then forms the first stage input), the right-hand sides, and rc_ode_step
// (error norm, accept or reject, the next step size, then the next attempt's step clipped at the next
// stop and its stage coefficients).  The host replays attempts without reading anything back until
// `done`; attempts replayed after the end do nothing but the (idempotent) last commit.
struct RcOdeCtl {
    double t, h, he, tn;
    long long attempts, accepted, rejected, maxiters;
    int si, node, nstops, npts;
    int done, active, commit, sample, last_rejected, clipped, status, pad;
};
// the persistent integration (rc_quartic_ode): operands, exchange buffers and the controller at entry
struct RcOdeRun {
    RcQuarticArgs g;           // the right-hand side's operands (U, F, epilogue fields unused)
    double* ybuf;              // [N] the state at entry, the end state on exit
    double* u0;                // [N] stage inputs exchanged between workgroups (two buffers)
    double* u1;
    double* slots;             // [nwg][2] per-tile error sums
    uint32_t* flags;           // [nwg] per-workgroup event counters, zero at entry
    double* Sf;                // samples, as OdeWork
    double* Sa;
    long long N, half;
    int nwg, ntrack, nranges, pad;
    const int* ranges;
    const double* stops;
    const int* stop_node;
    const double* cheb;
    RcOdeCtl c0;               // the controller at entry (the first attempt is set up in the kernel)
    double jac0, jt0, atol, rtol, cnt;
    long long wait_ticks;      // budget of one wait (constant-rate clock ticks)
    uint32_t* abort_flag;
    double* out;               // [8] status, attempts, accepted, rejected; workgroup 0's clock ticks in
                               // the products, the exchanges, the controller and the whole launch
};
__global__ void rc_quartic_ode(const RcOdeRun* pp);
#define RC_QUARTIC_ODE_MAX_WG 256   // workgroups one exchange round of rc_quartic_ode polls
hipError_t rc_quartic_ode_capacity(int* blocks);
hipError_t rc_launch_quartic_ode(dim3 grid, const RcOdeRun* dparams, hipStream_t stream);
__global__ void rc_ode_step(RcOdeCtl* c, int finish, const double* partial, const int* ranges, int nranges,
                            int ntrack, double cnt, const double* stops, const int* stop_node, const double* cheb,
                            double* cf);
__global__ void rc_ode_stage(const RcOdeCtl* c, long long N, long long half, double a10, const double* hp,
                             const double* ynew, double* y, const double* K12, double* K0, double* Sf, double* Sa,
                             double* U1);
__global__ void rc_trace_prod_batched(int Dp, const double* A, long long sa, const double* B, long long sb,
                                      double* out);
__global__ void rc_axpy_c(int n, double c_re, double c_im, const double* Z, double* X);
__global__ void rc_sub_scaled_n(int n, const double* t, const double* Z, double* X, long long sx);
__global__ void rc_add_identity_n(int Dp, int D, const double* t, double scale, double* X, long long sx);
__global__ void rc_norms_n(int n, const double* Xl, const double* X, long long sx, double* out, long long so);
__global__ void rc_poisson_scale_n(long long N, int Dp, int D, int adjoint, const double* s, int k0, const double* mu,
                                   const double* Y, long long sy, double* out, long long so);
__global__ void rc_axpby_n(int n, double ar, double ai, const double* X, long long sx, const double* b, const double* Y,
                           long long sy, double* out, long long so);
__global__ void rc_axpy_c_n(int n, const double* c, const double* Z, long long sz, double* X, long long sx);
__global__ void rc_lindblad_n(int Dp, int adjoint, const double* Q, const double* Qh, const double* R, const double* Rh,
                              const double* X, long long sx, double* out, long long so);
__global__ void rc_trace_axpy(int n, int nj, int nodes, const double* tr, const double* wz, double scale,
                              const double* Z, double* X);
__global__ void rc_trace_axpy_n(int n, int nj, int nodes, const double* tr, const double* wz, double scale,
                                const double* Z, long long sz, double* X, long long sx);
