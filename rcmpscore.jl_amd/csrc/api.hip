// C-ABI of librcmps_hip.so (declared in include/rcmps.h): state preparation, fg_phi4 and the
// boundary Poisson solve on one MI355X.  All O(D^3)-per-iteration work runs on the GPU; the host
// holds the Krylov bookkeeping (Hessenberg columns, Givens rotations) and, once per state, the
// D x D spectral factorisation of Q that the no-jump preconditioner applies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rcmps.h"
#include "host_la.h"
#include "linalg.h"
#include "rcmps_tables.h"
#include "trajectory.h"

int rc_trajectory_smem_bytes(int nb, int split);
hipError_t rc_trajectory_capacity(int nb, int split, int* blocks);
hipError_t rc_launch_trajectory(int nb, int split, int nblocks, const TrajParams* dparams, hipStream_t stream);

using rc::cd;
using rc::Mat;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    double* d() const { return static_cast<double*>(p); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct Fail {
    int code;
};

constexpr int kLive[RC_NSYS] = {1, 2, 1, 2, 3, 4};     // live channels per system
constexpr int kHasScalar[RC_NSYS] = {1, 0, 1, 0, 1, 0};

// trajectory workspace of one evaluation (one state) of a launch
struct TrajBufs {
    DevBuf Y[RC_NSYS], P[RC_NSYS], rec[RC_NSYS], samples[RC_NSYS], E, fixed, flags, out;
    size_t flag_bytes = 0;
    void release() {
        for (int s = 0; s < RC_NSYS; ++s) {
            Y[s].release(); P[s].release(); rec[s].release(); samples[s].release();
        }
        E.release(); fixed.release(); flags.release(); out.release();
    }
};

inline void ck(hipError_t e) {
    if (e != hipSuccess) throw Fail{RCMPS_ERR_DEVICE};
}

double host_source(double t) {
    double a = std::fabs(t);
    if (a >= RC_SOURCE_TMAX) return 0.0;
    int seg = (int)(a * (RC_SOURCE_SEGMENTS / RC_SOURCE_TMAX));
    if (seg > RC_SOURCE_SEGMENTS - 1) seg = RC_SOURCE_SEGMENTS - 1;
    double z = 2.0 * a * (RC_SOURCE_SEGMENTS / RC_SOURCE_TMAX) - 2.0 * seg - 1.0;
    const double* c = RC_SOURCE_CHEB + seg * (RC_SOURCE_DEGREE + 1);
    double b1 = 0, b2 = 0;
    for (int j = RC_SOURCE_DEGREE; j > 0; --j) {
        double b0 = 2 * z * b1 - b2 + c[j];
        b2 = b1;
        b1 = b0;
    }
    return z * b1 - b2 + c[0];
}

// Gauss-Legendre on (0, sqrt(max_point)) mirrored (src/observables.jl:65-75): Newton on P_n.
void quadrature(int n, double max_point, std::vector<double>& pts, std::vector<double>& wts) {
    std::vector<double> x(n), w(n);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) <= 4e-16 * std::fabs(z)) break;
        }
        double p0 = 1, p1 = z;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1 - z * z) * dp * dp);
    }
    const double a = std::sqrt(max_point);
    pts.assign(2 * n, 0.0);
    wts.assign(2 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double p = a / 2 * (x[i] + 1), wt = a / 2 * w[i];
        pts[n + i] = p;
        wts[n + i] = wt;
        pts[n - 1 - i] = -p;
        wts[n - 1 - i] = wt;
    }
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

struct rcmps_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // trajectory workspaces: per evaluation of a launch, and the launch description
    std::vector<std::unique_ptr<TrajBufs>> tb;
    DevBuf traj_params, traj_tasks, traj_sys, traj_sched, traj_points, abort_word;
    // pullback / misc workspaces
    DevBuf nodes_w, nodes_j, tmpM, partial, scratch, small;
    // krylov workspaces
    DevBuf kry_V, kry_w, kry_t1, kry_t2, kry_x, kry_b, kry_z, kry_h, kry_tmp;
    int last_D = 0, last_Dp = 0, last_npts = 0;
    bool last_adjoint = false;   // the last evaluation integrated the adjoint systems
    // quadrature cache: nodes (plus endpoint) and the pullback weights on the device
    int q_n = 0;
    double q_max_point = -1.0;
    std::vector<double> q_pts;
    hipEvent_t ev[8];
    double clock_mhz = 100.0;   // constant-rate device clock
};

struct rcmps_state {
    int D = 0, Dp = 0;
    DevBuf R, K, Q, Qh, Rh, C, Ch, rho, V, Vinv, lam, ident;
    double shift = 0.0;
    std::vector<cd> rho_host;
    double metric_reg_cached = -1.0;
    DevBuf Minv;   // (rho + reg I)^{-1}
};

namespace {

// ------------------------------------------------------------------------------------------
// device helpers on Dp x Dp matrices
// ------------------------------------------------------------------------------------------
struct Dev {
    rcmps_context* ctx;
    int Dp, D;
    int n() const { return Dp * Dp; }
    dim3 eb() const { return dim3(ceil_div(n(), 256)); }

    void gemm(int opA, const double* A, int opB, const double* B, double* C, double alpha = 1.0,
              double beta = 0.0) {
        GemmArgs g{};
        g.Dp = Dp; g.D = D; g.nbatch = 1; g.opA = opA; g.opB = opB;
        g.A = A; g.B = B; g.C = C;
        g.beta = beta;
        g.alpha_scale = alpha;
        hipLaunchKernelGGL(rc_zgemm_batched, dim3((Dp / 16) * (Dp / 16), 1), dim3(256), 0, ctx->stream, g);
        ck(hipGetLastError());
    }
    void batched(GemmArgs g, int nbatch) {
        g.Dp = Dp; g.D = D; g.nbatch = nbatch;
        hipLaunchKernelGGL(rc_zgemm_batched, dim3((Dp / 16) * (Dp / 16), nbatch), dim3(256), 0, ctx->stream, g);
        ck(hipGetLastError());
    }
    // C = beta C + scale * sum_b alpha_b op(A_b) op(B_b)
    void reduce(GemmArgs g, int nbatch, double* C, double scale, double beta) {
        g.Dp = Dp; g.D = D; g.nbatch = nbatch;
        const int nt = (Dp / 16) * (Dp / 16);
        int chunks = std::max(1, std::min(nbatch, 256 / nt * 2));
        int chunk = ceil_div(nbatch, chunks);
        chunks = ceil_div(nbatch, chunk);
        ck(ctx->partial.ensure((size_t)chunks * n() * 16));
        hipLaunchKernelGGL(rc_zgemm_reduce_partial, dim3(nt, chunks), dim3(256), 0, ctx->stream, g,
                           ctx->partial.d(), chunk);
        ck(hipGetLastError());
        hipLaunchKernelGGL(rc_sum_partials, eb(), dim3(256), 0, ctx->stream, ctx->partial.d(), chunks, n(), scale,
                           beta, C);
        ck(hipGetLastError());
    }
    void axpby(cd a, const double* X, cd b, const double* Y, double* out) {
        hipLaunchKernelGGL(rc_axpby, eb(), dim3(256), 0, ctx->stream, n(), a.real(), a.imag(), X, b.real(),
                           b.imag(), Y, out);
        ck(hipGetLastError());
    }
    void herm(int mode, double scale, const double* X, double* out) {
        hipLaunchKernelGGL(rc_herm, eb(), dim3(256), 0, ctx->stream, Dp, mode, scale, X, out);
        ck(hipGetLastError());
    }
    // out2 (device, 2 doubles) = tr(A B)
    void trace_prod(const double* A, const double* B, double* out2) {
        hipLaunchKernelGGL(rc_trace_prod, dim3(1), dim3(256), 0, ctx->stream, Dp, A, B, out2);
        ck(hipGetLastError());
    }
    void add_identity(const double* c2, double scale, double* X) {
        hipLaunchKernelGGL(rc_add_identity, dim3(ceil_div(D, 256)), dim3(256), 0, ctx->stream, Dp, D, c2, scale, X);
        ck(hipGetLastError());
    }
    void copy(const double* src, double* dst) {
        ck(hipMemcpyAsync(dst, src, (size_t)n() * 16, hipMemcpyDeviceToDevice, ctx->stream));
    }
    // X -= (t / den) Z with t = the device complex at t2
    void sub_scaled(const double* t2, cd den, const double* Z, double* X) {
        hipLaunchKernelGGL(rc_sub_scaled, eb(), dim3(256), 0, ctx->stream, n(), t2, den.real(), den.imag(), Z, X);
        ck(hipGetLastError());
    }
    // Y = -Y + scale * c I with c the device complex at c2
    void neg_add_identity(const double* c2, double scale, double* Y) {
        hipLaunchKernelGGL(rc_neg_add_identity, eb(), dim3(256), 0, ctx->stream, Dp, D, c2, scale, Y);
        ck(hipGetLastError());
    }
    cd fetch2(const double* d2) {
        double h[2];
        ck(hipMemcpyAsync(h, d2, 16, hipMemcpyDeviceToHost, ctx->stream));
        ck(hipStreamSynchronize(ctx->stream));
        return cd(h[0], h[1]);
    }
    double norm(const double* x) {
        double* out = ctx->small.d() + 32;
        hipLaunchKernelGGL(rc_dots, dim3(1), dim3(256), 0, ctx->stream, x, 0LL, n(), x, out);
        ck(hipGetLastError());
        return std::sqrt(std::max(0.0, fetch2(out).real()));
    }
};

void pad_in(rcmps_context* ctx, int D, int Dp, const double* src, double* dst) {
    ck(hipMemsetAsync(dst, 0, (size_t)Dp * Dp * 16, ctx->stream));
    ck(hipMemcpy2DAsync(dst, (size_t)Dp * 16, src, (size_t)D * 16, (size_t)D * 16, D, hipMemcpyDeviceToDevice,
                        ctx->stream));
}

void pad_out(rcmps_context* ctx, int D, int Dp, const double* src, double* dst) {
    ck(hipMemcpy2DAsync(dst, (size_t)D * 16, src, (size_t)Dp * 16, (size_t)D * 16, D, hipMemcpyDeviceToDevice,
                        ctx->stream));
}

Mat download(rcmps_context* ctx, int D, int Dp, const double* src) {
    std::vector<double> h((size_t)Dp * Dp * 2);
    ck(hipMemcpyAsync(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    Mat m((size_t)D * D);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) m[(size_t)i * D + j] = cd(h[2 * ((size_t)i * Dp + j)], h[2 * ((size_t)i * Dp + j) + 1]);
    return m;
}

void upload(rcmps_context* ctx, int D, int Dp, const Mat& m, double* dst) {
    std::vector<double> h((size_t)Dp * Dp * 2, 0.0);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            h[2 * ((size_t)i * Dp + j)] = m[(size_t)i * D + j].real();
            h[2 * ((size_t)i * Dp + j) + 1] = m[(size_t)i * D + j].imag();
        }
    ck(hipMemcpyAsync(dst, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
}

double fro(const Mat& a) {
    double s = 0;
    for (auto& z : a) s += std::norm(z);
    return std::sqrt(s);
}

// ------------------------------------------------------------------------------------------
// no-jump inverse in the eigenbasis of Q:  forward  X = P(Y) solves (Q - s/2) X + X (Q - s/2)^+ = -Y,
// adjoint X = P*(Y) solves (Q - s/2)^+ X + X (Q - s/2) = -Y  (src/transfer.jl:50-68).  Exact when Q
// is diagonalisable; used only as a right preconditioner, so the Krylov solution is exact regardless.
// ------------------------------------------------------------------------------------------
void nojump_apply(Dev& dv, const rcmps_state* st, const double* Y, double* out, bool adjoint, double* t1,
                  double* t2) {
    if (!adjoint) {
        dv.gemm(0, st->Vinv.d(), 0, Y, t1);
        dv.gemm(0, t1, 1, st->Vinv.d(), t2);
        hipLaunchKernelGGL(rc_sylvester_diag, dv.eb(), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, 0, st->shift,
                           st->lam.d(), t2, t1);
        ck(hipGetLastError());
        dv.gemm(0, st->V.d(), 0, t1, t2);
        dv.gemm(0, t2, 1, st->V.d(), out);
    } else {
        dv.gemm(1, st->V.d(), 0, Y, t1);
        dv.gemm(0, t1, 0, st->V.d(), t2);
        hipLaunchKernelGGL(rc_sylvester_diag, dv.eb(), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, 1, st->shift,
                           st->lam.d(), t2, t1);
        ck(hipGetLastError());
        dv.gemm(1, st->Vinv.d(), 0, t1, t2);
        dv.gemm(0, t2, 0, st->Vinv.d(), out);
    }
}

// L(X) = Q X + X Q^+ + R X R^+ ; L*(X) = Q^+ X + X Q + R^+ X R
void lindblad(Dev& dv, const rcmps_state* st, const double* X, double* out, bool adjoint, double* t1) {
    if (!adjoint) {
        dv.gemm(0, st->Q.d(), 0, X, out);
        dv.gemm(0, X, 0, st->Qh.d(), out, 1.0, 1.0);
        dv.gemm(0, st->R.d(), 0, X, t1);
        dv.gemm(0, t1, 0, st->Rh.d(), out, 1.0, 1.0);
    } else {
        dv.gemm(0, st->Qh.d(), 0, X, out);
        dv.gemm(0, X, 0, st->Q.d(), out, 1.0, 1.0);
        dv.gemm(0, X, 0, st->R.d(), t1);
        dv.gemm(0, st->Rh.d(), 0, t1, out, 1.0, 1.0);
    }
}

// Restarted GMRES(m) with classical Gram-Schmidt twice; vectors on the device, Hessenberg on the
// host.  Solves A x = b; returns true when the true residual is <= tol_abs.
bool gmres(Dev& dv, const std::function<void(const double*, double*)>& A, const double* b, double* x,
           double tol_abs, int m, int maxiter, int& matvecs, double& resid) {
    rcmps_context* ctx = dv.ctx;
    const int n = dv.n();
    ck(ctx->kry_V.ensure((size_t)(m + 1) * n * 16));
    ck(ctx->kry_w.ensure((size_t)n * 16));
    ck(ctx->kry_h.ensure((size_t)(4 * (m + 1) + 2) * 8 + 64));
    double* V = ctx->kry_V.d();
    double* w = ctx->kry_w.d();
    double* hdev = ctx->kry_h.d();
    ck(hipMemsetAsync(x, 0, (size_t)n * 16, ctx->stream));
    std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), sn(m), yk(m), hv(2 * (m + 1) + 1);
    std::vector<double> cs(m);
    matvecs = 0;
    bool converged = false;
    auto residual_into_w = [&]() {
        A(x, w);
        ++matvecs;
        dv.axpby(cd(1, 0), b, cd(-1, 0), w, w);
        return dv.norm(w);
    };
    double beta = residual_into_w();
    resid = beta;
    double beta_prev = INFINITY;
    for (int outer = 0; outer < maxiter && !converged; ++outer) {
        if (beta <= tol_abs) { converged = true; break; }
        if (beta > 0.9 * beta_prev) break;   // a restart cycle that gains under 10%: stagnation
        beta_prev = beta;
        dv.axpby(cd(1.0 / beta, 0), w, cd(0, 0), nullptr, V);
        std::fill(H.begin(), H.end(), cd(0, 0));
        std::fill(g.begin(), g.end(), cd(0, 0));
        g[0] = beta;
        int k = 0;
        for (; k < m; ++k) {
            A(V + (size_t)k * n * 2, w);
            ++matvecs;
            // classical Gram-Schmidt twice on the device: h1, h2 (k+1 complex each), then <w,w>
            double* hd1 = hdev;
            double* hd2 = hdev + 2 * (m + 1);
            double* hnn = hdev + 4 * (m + 1);
            hipLaunchKernelGGL(rc_dots, dim3(k + 1), dim3(256), 0, ctx->stream, V, (long long)n, n, w, hd1);
            hipLaunchKernelGGL(rc_subtract_combo, dv.eb(), dim3(256), 0, ctx->stream, V, (long long)n, k + 1, hd1, n, w);
            hipLaunchKernelGGL(rc_dots, dim3(k + 1), dim3(256), 0, ctx->stream, V, (long long)n, n, w, hd2);
            hipLaunchKernelGGL(rc_subtract_combo, dv.eb(), dim3(256), 0, ctx->stream, V, (long long)n, k + 1, hd2, n, w);
            hipLaunchKernelGGL(rc_dots, dim3(1), dim3(256), 0, ctx->stream, w, 0LL, n, w, hnn);
            hipLaunchKernelGGL(rc_normalize, dv.eb(), dim3(256), 0, ctx->stream, n, w, hnn, V + (size_t)(k + 1) * n * 2);
            ck(hipGetLastError());
            ck(hipMemcpyAsync(hv.data(), hdev, (size_t)(4 * (m + 1) + 2) * 8, hipMemcpyDeviceToHost, ctx->stream));
            ck(hipStreamSynchronize(ctx->stream));
            for (int i = 0; i <= k; ++i)
                H[(size_t)i * m + k] += hv[i] + hv[(m + 1) + i];
            double hn = std::sqrt(std::max(0.0, hv[2 * (m + 1)].real()));
            H[(size_t)(k + 1) * m + k] = hn;
            for (int i = 0; i < k; ++i) {
                cd h1 = H[(size_t)i * m + k], h2 = H[(size_t)(i + 1) * m + k];
                H[(size_t)i * m + k] = cs[i] * h1 + sn[i] * h2;
                H[(size_t)(i + 1) * m + k] = -std::conj(sn[i]) * h1 + cs[i] * h2;
            }
            {
                cd xx = H[(size_t)k * m + k], yy = H[(size_t)(k + 1) * m + k];
                double ax = std::abs(xx), ay = std::abs(yy);
                if (ay == 0) { cs[k] = 1; sn[k] = 0; }
                else if (ax == 0) { cs[k] = 0; sn[k] = std::conj(yy) / ay; }
                else { double r = std::hypot(ax, ay); cs[k] = ax / r; sn[k] = (xx / ax) * std::conj(yy) / r; }
            }
            cd h1 = H[(size_t)k * m + k], h2 = H[(size_t)(k + 1) * m + k];
            H[(size_t)k * m + k] = cs[k] * h1 + sn[k] * h2;
            H[(size_t)(k + 1) * m + k] = 0;
            cd g1 = g[k], g2 = g[k + 1];
            g[k] = cs[k] * g1 + sn[k] * g2;
            g[k + 1] = -std::conj(sn[k]) * g1 + cs[k] * g2;
            if (std::abs(g[k + 1]) <= tol_abs || hn == 0) { ++k; break; }
        }
        const int kk = std::min(k, m);
        for (int i = kk - 1; i >= 0; --i) {
            cd s = g[i];
            for (int l = i + 1; l < kk; ++l) s -= H[(size_t)i * m + l] * yk[l];
            yk[i] = s / H[(size_t)i * m + i];
        }
        ck(hipMemcpyAsync(hdev, yk.data(), (size_t)kk * 16, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(rc_add_combo, dv.eb(), dim3(256), 0, ctx->stream, V, (long long)n, kk, hdev, n, x);
        ck(hipGetLastError());
        beta = residual_into_w();
        resid = beta;
        if (beta <= tol_abs) converged = true;
    }
    return converged;
}

void state_spectral(rcmps_context* ctx, rcmps_state* st) {
    const int D = st->D, Dp = st->Dp;
    Mat q = download(ctx, D, Dp, st->Q.d());
    Mat r = download(ctx, D, Dp, st->R.d());
    Mat T = q, Z;
    if (!rc::schur(D, T, Z)) throw Fail{RCMPS_ERR_SOLVER};
    Mat W, Winv;
    rc::triangular_eigenvectors(D, T, W);
    if (!rc::upper_inverse(D, W, Winv)) throw Fail{RCMPS_ERR_SOLVER};
    Mat V, Vinv;
    rc::matmul(D, Z, W, V);
    rc::matmul(D, Winv, rc::adjoint(D, Z), Vinv);
    double nq = rc::opnorm2(D, q), nr = rc::opnorm2(D, r);
    double scale = std::max(1.0, std::max(nq, nr * nr));
    double maxre = -INFINITY;
    std::vector<double> lam((size_t)Dp * 2, 0.0);
    for (int i = 0; i < D; ++i) {
        cd l = T[(size_t)i * D + i];
        maxre = std::max(maxre, l.real());
        lam[2 * i] = l.real();
        lam[2 * i + 1] = l.imag();
    }
    st->shift = maxre >= -100 * 2.220446049250313e-16 * scale ? scale : 0.0;
    ck(st->V.ensure((size_t)Dp * Dp * 16));
    ck(st->Vinv.ensure((size_t)Dp * Dp * 16));
    ck(st->lam.ensure((size_t)Dp * 16));
    upload(ctx, D, Dp, V, st->V.d());
    upload(ctx, D, Dp, Vinv, st->Vinv.d());
    ck(hipMemcpyAsync(st->lam.d(), lam.data(), lam.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
}

// checked_density (src/transfer.jl:33-47) on the host copy; returns the normalised density
bool checked_density(int D, const Mat& q, const Mat& r, Mat& rho, double tol) {
    for (auto& z : rho)
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) return false;
    cd z = 0;
    for (int i = 0; i < D; ++i) z += rho[(size_t)i * D + i];
    if (!(std::abs(z) > 2.220446049250313e-16 * fro(rho))) return false;
    for (auto& x : rho) x /= z;
    double herm = 0;
    Mat x((size_t)D * D);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            cd a = rho[(size_t)i * D + j], bt = std::conj(rho[(size_t)j * D + i]);
            herm += std::norm(a - bt);
            x[(size_t)i * D + j] = (a + bt) / 2.0;
        }
    if (!(std::sqrt(herm) <= 100 * tol)) return false;
    rho = x;
    Mat t1, t2, l((size_t)D * D, cd(0, 0));
    rc::matmul(D, q, rho, t1);
    rc::matmul(D, rho, rc::adjoint(D, q), t2);
    Mat t3, t4;
    rc::matmul(D, r, rho, t3);
    rc::matmul(D, t3, rc::adjoint(D, r), t4);
    for (size_t i = 0; i < l.size(); ++i) l[i] = t1[i] + t2[i] + t4[i];
    double scale = std::max(1.0, 2 * rc::opnorm2(D, q) + std::pow(rc::opnorm2(D, r), 2));
    if (!(fro(l) <= tol * scale)) return false;
    Mat T = rho, Z;
    if (!rc::schur(D, T, Z)) return false;
    double mn = INFINITY;
    for (int i = 0; i < D; ++i) mn = std::min(mn, T[(size_t)i * D + i].real());
    return mn >= -10 * tol;
}

int guard(const std::function<void()>& f) {
    try {
        f();
        return RCMPS_OK;
    } catch (const Fail& e) {
        return e.code;
    }
}

}  // namespace

// ==========================================================================================
// C ABI
// ==========================================================================================
extern "C" {

const char* rcmps_status_string(int s) {
    switch (s) {
        case RCMPS_OK: return "success";
        case RCMPS_ERR_MAX_STEPS: return "max_steps";
        case RCMPS_ERR_STEP_UNDERFLOW: return "step_underflow";
        case RCMPS_ERR_NONFINITE: return "nonfinite_error";
        case RCMPS_ERR_MISSING_SAMPLES: return "missing_samples";
        case RCMPS_ERR_WAIT_LIMIT: return "wait_limit";
        case RCMPS_ERR_SOLVER: return "solver_failure";
        case RCMPS_ERR_ARGUMENT: return "invalid_argument";
        case RCMPS_ERR_DEVICE: return "device_error";
        case RCMPS_ERR_UNSUPPORTED: return "unsupported";
        default: return "unknown";
    }
}

int rcmps_max_bond_dimension(void) { return 64; }

int rcmps_context_create(int device, rcmps_context** out) {
    if (!out) return RCMPS_ERR_ARGUMENT;
    auto* ctx = new rcmps_context();
    ctx->device = device;
    int rc = guard([&] {
        ck(hipSetDevice(device));
        ck(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        for (auto& e : ctx->ev) ck(hipEventCreate(&e));
        ck(ctx->small.ensure(4096));
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            ctx->clock_mhz = khz / 1000.0;
        ck(ctx->abort_word.ensure(256));
    });
    if (rc) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return RCMPS_OK;
}

void rcmps_context_destroy(rcmps_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf* all[] = {&ctx->traj_params, &ctx->traj_tasks, &ctx->traj_sys, &ctx->traj_sched,
                     &ctx->traj_points, &ctx->abort_word, &ctx->nodes_w, &ctx->nodes_j,
                     &ctx->tmpM, &ctx->partial, &ctx->scratch, &ctx->small, &ctx->kry_V, &ctx->kry_w,
                     &ctx->kry_t1, &ctx->kry_t2, &ctx->kry_x, &ctx->kry_b, &ctx->kry_z, &ctx->kry_h,
                     &ctx->kry_tmp};
    for (auto* b : all) b->release();
    for (auto& b : ctx->tb) b->release();
    for (auto& e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void rcmps_state_destroy(rcmps_state* st) {
    if (!st) return;
    DevBuf* all[] = {&st->R, &st->K, &st->Q, &st->Qh, &st->Rh, &st->C, &st->Ch, &st->rho,
                     &st->V, &st->Vinv, &st->lam, &st->ident, &st->Minv};
    for (auto* b : all) b->release();
    delete st;
}

int rcmps_state_create(rcmps_context* ctx, int D, const double* R, const double* K, const double* rho,
                       double steady_tol, rcmps_state** out, rcmps_stats* stats) {
    if (!ctx || !out || !R || !K || D < 1 || !(steady_tol > 0)) return RCMPS_ERR_ARGUMENT;
    auto* st = new rcmps_state();
    st->D = D;
    st->Dp = ((D + 15) / 16) * 16;
    const int Dp = st->Dp;
    const size_t mb = (size_t)Dp * Dp * 16;
    int rc = guard([&] {
        ck(hipSetDevice(ctx->device));
        DevBuf* mats[] = {&st->R, &st->K, &st->Q, &st->Qh, &st->Rh, &st->C, &st->Ch, &st->rho, &st->ident};
        for (auto* b : mats) ck(b->ensure(mb));
        ck(ctx->kry_t1.ensure(mb));
        ck(ctx->kry_t2.ensure(mb));
        Dev dv{ctx, Dp, D};
        pad_in(ctx, D, Dp, R, st->R.d());
        pad_in(ctx, D, Dp, K, st->K.d());
        // validation: finite, K Hermitian (src/state.jl:41-44)
        Mat kh = download(ctx, D, Dp, st->K.d());
        Mat rh = download(ctx, D, Dp, st->R.d());
        for (auto& z : kh) if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) throw Fail{RCMPS_ERR_ARGUMENT};
        for (auto& z : rh) if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) throw Fail{RCMPS_ERR_ARGUMENT};
        double dk = 0;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) dk += std::norm(kh[(size_t)i * D + j] - std::conj(kh[(size_t)j * D + i]));
        if (!(std::sqrt(dk) <= 1e-12 * std::max(1.0, fro(kh)))) throw Fail{RCMPS_ERR_ARGUMENT};
        dv.herm(0, 1.0, st->K.d(), ctx->kry_t1.d());
        dv.copy(ctx->kry_t1.d(), st->K.d());
        // Q = -i K - R^+ R / 2
        dv.gemm(1, st->R.d(), 0, st->R.d(), ctx->kry_t1.d());
        dv.axpby(cd(0, -1), st->K.d(), cd(-0.5, 0), ctx->kry_t1.d(), st->Q.d());
        dv.herm(1, 1.0, st->Q.d(), st->Qh.d());
        dv.herm(1, 1.0, st->R.d(), st->Rh.d());
        dv.gemm(0, st->Q.d(), 0, st->R.d(), st->C.d());
        dv.gemm(0, st->R.d(), 0, st->Q.d(), st->C.d(), -1.0, 1.0);
        dv.herm(1, 1.0, st->C.d(), st->Ch.d());
        // identity on the D x D block
        {
            std::vector<double> h((size_t)Dp * Dp * 2, 0.0);
            for (int i = 0; i < D; ++i) h[2 * ((size_t)i * Dp + i)] = 1.0;
            ck(hipMemcpyAsync(st->ident.d(), h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        }
        state_spectral(ctx, st);
        Mat qh = download(ctx, D, Dp, st->Q.d());
        Mat rho_h;
        if (rho) {
            pad_in(ctx, D, Dp, rho, st->rho.d());
            rho_h = download(ctx, D, Dp, st->rho.d());
        } else if (D == 1) {
            rho_h = Mat(1, cd(1, 0));
        } else {
            // (-L + |I><I|/D) x = I/D with the Sherman-Morrison right preconditioner built on the
            // no-jump inverse (gpu/src/rcmps_gpu/optimizer/runtime/numerics.py:74-100)
            double* t1 = ctx->kry_t1.d();
            double* t2 = ctx->kry_t2.d();
            ck(ctx->kry_z.ensure(mb));
            ck(ctx->kry_b.ensure(mb));
            ck(ctx->kry_x.ensure(mb));
            ck(ctx->scratch.ensure(4 * mb));
            double* z = ctx->kry_z.d();
            double* tmp = ctx->scratch.d();
            double* tmp2 = tmp + (size_t)Dp * Dp * 2;
            double* tr2 = ctx->small.d();
            nojump_apply(dv, st, st->ident.d(), z, false, t1, t2);
            dv.trace_prod(st->ident.d(), z, tr2);
            cd den = 1.0 + dv.fetch2(tr2) / (double)D;
            auto pre = [&](const double* y, double* xo) {
                nojump_apply(dv, st, y, xo, false, t1, t2);
                dv.trace_prod(st->ident.d(), xo, tr2);
                dv.sub_scaled(tr2, (double)D * den, z, xo);
            };
            auto op = [&](const double* xin, double* yo) {
                lindblad(dv, st, xin, yo, false, t1);
                dv.trace_prod(st->ident.d(), xin, tr2);
                dv.neg_add_identity(tr2, 1.0 / D, yo);
            };
            dv.axpby(cd(1.0 / D, 0), st->ident.d(), cd(0, 0), nullptr, ctx->kry_b.d());
            int mv = 0;
            double res = 0;
            auto A = [&](const double* v, double* w) { pre(v, tmp2); op(tmp2, w); };
            const int m = std::min(D * D, 40);
            bool okg = gmres(dv, A, ctx->kry_b.d(), ctx->kry_x.d(), steady_tol, m, 200, mv, res);
            pre(ctx->kry_x.d(), st->rho.d());
            if (stats) { stats->krylov_matvecs = mv; stats->krylov_residual = res; }
            if (!okg) throw Fail{RCMPS_ERR_SOLVER};
            rho_h = download(ctx, D, Dp, st->rho.d());
        }
        if (!checked_density(D, qh, rh, rho_h, steady_tol)) throw Fail{RCMPS_ERR_SOLVER};
        upload(ctx, D, Dp, rho_h, st->rho.d());
        st->rho_host = rho_h;
        ck(hipStreamSynchronize(ctx->stream));
    });
    if (rc) {
        rcmps_state_destroy(st);
        return rc;
    }
    *out = st;
    return RCMPS_OK;
}

int rcmps_state_get(rcmps_context* ctx, const rcmps_state* st, int which, double* dst) {
    if (!ctx || !st || !dst || (which != 0 && which != 1)) return RCMPS_ERR_ARGUMENT;
    return guard([&] {
        pad_out(ctx, st->D, st->Dp, which == 0 ? st->Q.d() : st->rho.d(), dst);
        ck(hipStreamSynchronize(ctx->stream));
    });
}

namespace {

// boundary Poisson solve on padded device matrices: h solves (-L* + |I><rho|) h = a - tr(rho a) I
// GMRES aims at `tol`; the solution is accepted when the true residual is within 20 accept_tol
// (relative to max(1, |b|)), the bar of src/transfer.jl:140-152.
void boundary_solve(rcmps_context* ctx, const rcmps_state* st, const double* a, double tol, double accept_tol,
                    double* h, rcmps_stats* stats) {
    const int D = st->D, Dp = st->Dp;
    const size_t mb = (size_t)Dp * Dp * 16;
    Dev dv{ctx, Dp, D};
    ck(ctx->kry_t1.ensure(mb));
    ck(ctx->kry_t2.ensure(mb));
    ck(ctx->kry_z.ensure(mb));
    ck(ctx->kry_b.ensure(mb));
    ck(ctx->kry_x.ensure(mb));
    ck(ctx->kry_tmp.ensure(2 * mb));
    double* t1 = ctx->kry_t1.d();
    double* t2 = ctx->kry_t2.d();
    double* z = ctx->kry_z.d();
    double* b = ctx->kry_b.d();
    double* tmp = ctx->kry_tmp.d();
    double* tmp2 = tmp + (size_t)Dp * Dp * 2;
    double* tr2 = ctx->small.d();
    if (D == 1) {
        ck(hipMemsetAsync(h, 0, mb, ctx->stream));
        return;
    }
    dv.copy(a, b);
    dv.trace_prod(st->rho.d(), a, tr2);
    dv.add_identity(tr2, -1.0, b);
    const double nb = dv.norm(b);
    nojump_apply(dv, st, st->ident.d(), z, true, t1, t2);
    dv.trace_prod(st->rho.d(), z, tr2);
    const cd den = 1.0 + dv.fetch2(tr2);
    auto pre = [&](const double* y, double* xo) {
        nojump_apply(dv, st, y, xo, true, t1, t2);
        dv.trace_prod(st->rho.d(), xo, tr2);
        dv.sub_scaled(tr2, den, z, xo);
    };
    auto op = [&](const double* xin, double* yo) {
        lindblad(dv, st, xin, yo, true, t1);
        dv.trace_prod(st->rho.d(), xin, tr2);
        dv.neg_add_identity(tr2, 1.0, yo);
    };
    auto A = [&](const double* v, double* w) { pre(v, tmp2); op(tmp2, w); };
    const double tol_abs = tol * std::max(1.0, nb);
    int mv = 0;
    double res = 0;
    const int m = std::min(D * D, 40);
    bool ok = gmres(dv, A, b, ctx->kry_x.d(), tol_abs, m, 200, mv, res);
    pre(ctx->kry_x.d(), h);
    // true residual of the physical equation (src/transfer.jl:148-152)
    op(h, tmp);
    dv.axpby(cd(1, 0), tmp, cd(-1, 0), b, tmp);
    const double true_res = dv.norm(tmp);
    if (stats) { stats->krylov_matvecs = mv; stats->krylov_residual = true_res; }
    (void)ok;
    if (!(true_res <= 20 * std::max(tol, accept_tol) * std::max(1.0, nb))) throw Fail{RCMPS_ERR_SOLVER};
}

}  // namespace

int rcmps_boundary_environment(rcmps_context* ctx, const rcmps_state* st, const double* a, double tol,
                               double* h, rcmps_stats* stats) {
    if (!ctx || !st || !a || !h || !(tol > 0)) return RCMPS_ERR_ARGUMENT;
    return guard([&] {
        const size_t mb = (size_t)st->Dp * st->Dp * 16;
        ck(ctx->tmpM.ensure(2 * mb));
        double* ap = ctx->tmpM.d();
        double* hp = ap + (size_t)st->Dp * st->Dp * 2;
        pad_in(ctx, st->D, st->Dp, a, ap);
        boundary_solve(ctx, st, ap, tol, tol, hp, stats);
        pad_out(ctx, st->D, st->Dp, hp, h);
        ck(hipStreamSynchronize(ctx->stream));
    });
}

namespace {

// One fg_phi4 evaluation of a batch: its state, coupling and outputs.
struct EvalReq {
    const rcmps_state* st = nullptr;
    double g = 0.0;
    double* energy = nullptr;                                    // host
    double *grad = nullptr, *raw_cov = nullptr, *raw_env = nullptr, *grad_R = nullptr, *grad_K = nullptr;
    rcmps_stats* S = nullptr;
    bool adjoint() const { return grad || raw_cov || raw_env || grad_R || grad_K; }
};


void ensure_quadrature(rcmps_context* ctx, const rcmps_integrator* ip) {
    if (ctx->q_n == ip->n_quad && ctx->q_max_point == ip->max_point) return;
    const int nn = 2 * ip->n_quad, npts = nn + 1;
    std::vector<double> pts, wts;
    quadrature(ip->n_quad, ip->max_point, pts, wts);
    pts.push_back(std::sqrt(ip->max_point));
    std::vector<double> wv(nn), jv(nn);
    for (int i = 0; i < nn; ++i) {
        wv[i] = wts[i] * 2 * std::fabs(pts[i]);
        jv[i] = wts[i] * host_source(pts[i]);
    }
    ck(ctx->traj_points.ensure((size_t)npts * 8));
    ck(ctx->nodes_w.ensure((size_t)nn * 8));
    ck(ctx->nodes_j.ensure((size_t)nn * 8));
    ck(hipMemcpyAsync(ctx->traj_points.d(), pts.data(), (size_t)npts * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(ctx->nodes_w.d(), wv.data(), (size_t)nn * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(ctx->nodes_j.d(), jv.data(), (size_t)nn * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    ctx->q_n = ip->n_quad;
    ctx->q_max_point = ip->max_point;
    ctx->q_pts = pts;
}

// Fixed insertion sources, the six systems and their tasks for one state (src/observables.jl:7-29:
// kinetic edges sqrt2 with op R or [Q,R], :phi^4: edges k -> k+1 with coefficient k).
void build_state(rcmps_context* ctx, TrajBufs& tb, const rcmps_state* st, bool adjoint, int npts,
                 std::vector<TrajSys>& sys, std::vector<TrajTask>& tasks) {
    const int D = st->D, Dp = st->Dp, NB = Dp / 16, NT = NB * NB;
    Dev dv{ctx, Dp, D};
    const size_t mb = (size_t)Dp * Dp * 16;
    const size_t mcount = (size_t)Dp * Dp * 2;
    ck(tb.fixed.ensure(6 * mb));
    double* fx = tb.fixed.d();
    double* fK1F = fx;
    double* fK1A = fx + mcount;
    double* fK2F = fx + 2 * mcount;
    double* fK2A = fx + 3 * mcount;
    double* fQF = fx + 4 * mcount;
    double* fQA = fx + 5 * mcount;
    const double s2 = std::sqrt(2.0);
    dv.gemm(0, st->R.d(), 0, st->rho.d(), fK1F, s2);                 // sqrt2 R rho
    dv.axpby(cd(s2, 0), st->Rh.d(), cd(0, 0), nullptr, fK1A);         // sqrt2 R^+
    dv.gemm(0, st->C.d(), 0, st->rho.d(), fK2F, s2);                 // sqrt2 C rho
    dv.axpby(cd(s2, 0), st->Ch.d(), cd(0, 0), nullptr, fK2A);         // sqrt2 C^+
    dv.gemm(0, st->R.d(), 0, st->rho.d(), fQF);                      // R rho + rho R^+
    dv.gemm(0, st->rho.d(), 0, st->Rh.d(), fQF, 1.0, 1.0);
    dv.axpby(cd(4, 0), st->R.d(), cd(4, 0), st->Rh.d(), fQA);         // 4 (R + R^+)

    size_t flag_words = 0;
    for (int s = 0; s < RC_NSYS; ++s) flag_words += (size_t)2 * kLive[s] * NT + NT + 64;
    ck(tb.flags.ensure(flag_words * 4));
    tb.flag_bytes = flag_words * 4;
    ck(tb.out.ensure(RC_NSYS * RC_OUT_STRIDE * 8));
    ck(tb.E.ensure((size_t)RC_NSTAGE_BUF * mb));
    uint32_t* fl = static_cast<uint32_t*>(tb.flags.p);
    const int sys_base = (int)sys.size();
    for (int s = 0; s < RC_NSYS; ++s) {
        TrajSys Sy{};
        Sy.nlive = kLive[s];
        Sy.has_scalar = kHasScalar[s];
        Sy.adj = s % 2;
        Sy.first_task = (int)tasks.size();
        Sy.ntasks = !adjoint && s % 2 ? 0 : kLive[s] * NT;
        Sy.D = D;
        Sy.n_elems = (double)kLive[s] * D * D + kHasScalar[s];
        Sy.Q = st->Q.d(); Sy.Qh = st->Qh.d(); Sy.R = st->R.d(); Sy.Rh = st->Rh.d(); Sy.C = st->C.d();
        Sy.flagY = fl; fl += kLive[s] * NT;
        Sy.flagP = fl; fl += kLive[s] * NT;
        Sy.flagE = fl; fl += NT;
        Sy.flagR = fl; fl += 64;
        Sy.out = tb.out.d() + s * RC_OUT_STRIDE;
        if (Sy.ntasks > 64) throw Fail{RCMPS_ERR_UNSUPPORTED};
        if (Sy.ntasks > 0) {
            if ((size_t)npts * kLive[s] * mb >= (size_t)1 << 31) throw Fail{RCMPS_ERR_UNSUPPORTED};
            ck(tb.Y[s].ensure((size_t)RC_NSTAGE_BUF * kLive[s] * mb));
            ck(tb.P[s].ensure((size_t)RC_NSTAGE_BUF * kLive[s] * mb));
            ck(tb.rec[s].ensure((size_t)2 * Sy.ntasks * RC_REC * 8));
            ck(tb.samples[s].ensure((size_t)npts * kLive[s] * mb));
            Sy.Y = tb.Y[s].d();
            Sy.P = tb.P[s].d();
            Sy.E = s == 3 ? tb.E.d() : nullptr;
            Sy.rec = tb.rec[s].d();
            Sy.samples = tb.samples[s].d();
        }
        const int term = s / 2, adj = s % 2;
        for (int ch = 0; ch < kLive[s] && Sy.ntasks > 0; ++ch)
            for (int b = 0; b < NB; ++b)
                for (int c = 0; c < NB; ++c) {
                    TrajTask t{};
                    t.sys = sys_base + s; t.ch = ch; t.b = b; t.c = c; t.adj = adj;
                    t.edge_src = -1;
                    if (term < 2) {
                        const double* O = term == 0 ? st->R.d() : st->C.d();
                        if (!adj) {
                            t.fixed = term == 0 ? fK1F : fK2F;
                            t.trace_kind = 1;
                            t.trace_op = O;
                            t.trace_coef = 2 * s2;
                        } else if (ch == 0) {
                            t.edge_src = 1;
                            t.edge_from_E = term == 1;
                            t.edge_coef = s2;
                        } else {
                            t.fixed = term == 0 ? fK1A : fK2A;
                            t.has_E = term == 1;
                        }
                    } else if (!adj) {
                        if (ch == 0) t.fixed = fQF;
                        else { t.edge_src = ch - 1; t.edge_coef = ch + 1; }
                        if (ch == 2) { t.trace_kind = 2; t.trace_coef = 8.0; }
                    } else {
                        if (ch < 3) { t.edge_src = ch + 1; t.edge_coef = ch + 1; }
                        else t.fixed = fQA;
                    }
                    tasks.push_back(t);
                }
        sys.push_back(Sy);
    }
}

// Integrate every system of the given evaluations (all of one padded size Dp) in one persistent
// launch; states are packed into rounds of at most the resident workgroup count.
void run_class(rcmps_context* ctx, int Dp, bool split, const rcmps_integrator* ip, const std::vector<int>& idx,
               std::vector<EvalReq>& reqs) {
    const int NB = Dp / 16, npts = 2 * ip->n_quad + 1;
    std::vector<TrajSys> sys;
    std::vector<TrajTask> tasks;
    while (ctx->tb.size() < idx.size()) ctx->tb.emplace_back(new TrajBufs());
    std::vector<int> state_first, state_ntasks;
    for (size_t j = 0; j < idx.size(); ++j) {
        const int first = (int)tasks.size();
        build_state(ctx, *ctx->tb[j], reqs[idx[j]].st, reqs[idx[j]].adjoint(), npts, sys, tasks);
        state_first.push_back(first);
        state_ntasks.push_back((int)tasks.size() - first);
    }
    int capacity = 0;
    ck(rc_trajectory_capacity(NB, split ? 1 : 0, &capacity));
    // gang schedule: a state's tasks share one round
    std::vector<std::vector<int>> rounds;
    int nblocks = 0;
    for (size_t j = 0; j < idx.size(); ++j) {
        if (state_ntasks[j] > capacity) throw Fail{RCMPS_ERR_UNSUPPORTED};
        if (rounds.empty() || (int)rounds.back().size() + state_ntasks[j] > capacity) rounds.emplace_back();
        for (int k = 0; k < state_ntasks[j]; ++k) rounds.back().push_back(state_first[j] + k);
        nblocks = std::max(nblocks, (int)rounds.back().size());
    }
    std::vector<int> sched((size_t)rounds.size() * nblocks, -1);
    for (size_t r = 0; r < rounds.size(); ++r)
        for (size_t w = 0; w < rounds[r].size(); ++w) sched[r * nblocks + w] = rounds[r][w];

    TrajParams P{};
    P.Dp = Dp;
    P.npoints = npts;
    P.nrounds = (int)rounds.size();
    P.a = std::sqrt(ip->max_point);
    P.atol = ip->atol;
    P.rtol = ip->rtol;
    P.h0 = ip->initial_step > 0 ? ip->initial_step : 0.01;
    P.max_steps = ip->maxiters;
    P.spin_limit = 20000000LL;
    P.points = ctx->traj_points.d();
    P.abort_flag = static_cast<uint32_t*>(ctx->abort_word.p);
    ck(ctx->traj_tasks.ensure(tasks.size() * sizeof(TrajTask)));
    ck(ctx->traj_sys.ensure(sys.size() * sizeof(TrajSys)));
    ck(ctx->traj_sched.ensure(sched.size() * sizeof(int)));
    ck(ctx->traj_params.ensure(sizeof(TrajParams) + 64 + sizeof(RC_SOURCE_CHEB)));
    double* cheb = reinterpret_cast<double*>(static_cast<char*>(ctx->traj_params.p) + sizeof(TrajParams) + 64);
    P.cheb = cheb;
    P.tasks = static_cast<const TrajTask*>(ctx->traj_tasks.p);
    P.sys = static_cast<const TrajSys*>(ctx->traj_sys.p);
    P.schedule = static_cast<const int*>(ctx->traj_sched.p);
    ck(hipMemcpyAsync(cheb, RC_SOURCE_CHEB, sizeof(RC_SOURCE_CHEB), hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(ctx->traj_tasks.p, tasks.data(), tasks.size() * sizeof(TrajTask), hipMemcpyHostToDevice,
                      ctx->stream));
    ck(hipMemcpyAsync(ctx->traj_sys.p, sys.data(), sys.size() * sizeof(TrajSys), hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(ctx->traj_sched.p, sched.data(), sched.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(ctx->traj_params.p, &P, sizeof(TrajParams), hipMemcpyHostToDevice, ctx->stream));
    for (size_t j = 0; j < idx.size(); ++j) {
        ck(hipMemsetAsync(ctx->tb[j]->flags.p, 0, ctx->tb[j]->flag_bytes, ctx->stream));
        ck(hipMemsetAsync(ctx->tb[j]->out.p, 0, RC_NSYS * RC_OUT_STRIDE * 8, ctx->stream));
    }
    ck(hipMemsetAsync(ctx->abort_word.p, 0, 256, ctx->stream));
    ck(hipEventRecord(ctx->ev[0], ctx->stream));
    ck(rc_launch_trajectory(NB, split ? 1 : 0, nblocks, static_cast<const TrajParams*>(ctx->traj_params.p),
                            ctx->stream));
    ck(hipEventRecord(ctx->ev[1], ctx->stream));
    std::vector<double> outs(idx.size() * RC_NSYS * RC_OUT_STRIDE);
    for (size_t j = 0; j < idx.size(); ++j)
        ck(hipMemcpyAsync(outs.data() + j * RC_NSYS * RC_OUT_STRIDE, ctx->tb[j]->out.d(), RC_NSYS * RC_OUT_STRIDE * 8,
                          hipMemcpyDeviceToHost, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    float tms = 0;
    ck(hipEventElapsedTime(&tms, ctx->ev[0], ctx->ev[1]));
    for (size_t j = 0; j < idx.size(); ++j) {
        EvalReq& r = reqs[idx[j]];
        rcmps_stats* S = r.S;
        S->trajectory_ms = tms;
        S->status = RCMPS_OK;
        S->failed_system = -1;
        for (int s = 0; s < RC_NSYS; ++s) {
            const double* o = outs.data() + (j * RC_NSYS + s) * RC_OUT_STRIDE;
            S->attempts[s] = (long long)o[RC_OUT_ATTEMPTS];
            S->accepted[s] = (long long)o[RC_OUT_ACCEPTED];
            S->rejected[s] = (long long)o[RC_OUT_REJECTED];
            S->dense_steps[s] = (long long)o[RC_OUT_DENSE];
            S->rhs_calls[s] = (long long)o[RC_OUT_RHS];
            for (int q = 0; q < RC_NTIMERS; ++q) S->phase_us[s][q] = o[RC_OUT_TIMERS + q] / ctx->clock_mhz;
            const int stt = (int)o[RC_OUT_STATUS];
            if (stt != RC_ST_OK && S->status == RCMPS_OK) { S->status = stt; S->failed_system = s; }
        }
        if (S->status == RCMPS_OK) {
            for (int term = 0; term < 3; ++term)
                S->term_values[term] = outs[(j * RC_NSYS + 2 * term) * RC_OUT_STRIDE + RC_OUT_SCALAR];
            const double coef[3] = {1.0, 1.0, r.g};
            double E = 0;
            for (int term = 0; term < 3; ++term) E += coef[term] * S->term_values[term];
            *r.energy = E;
        }
    }
    if (!idx.empty()) {
        const rcmps_state* st0 = reqs[idx[0]].st;
        ctx->last_D = st0->D; ctx->last_Dp = st0->Dp; ctx->last_npts = npts;
        ctx->last_adjoint = reqs[idx[0]].adjoint();
    }
}

// Pullback over the Gauss nodes, boundary correction and gradients of one evaluation whose
// trajectories are in tb (src/observables.jl:161-212).
void finish_state(rcmps_context* ctx, EvalReq& r, TrajBufs& tb, const rcmps_integrator* ip, double metric_reg,
                  int flags) {
    const rcmps_state* st = r.st;
    rcmps_stats* S = r.S;
    double *grad = r.grad, *raw_cov = r.raw_cov, *raw_env = r.raw_env, *grad_R = r.grad_R, *grad_K = r.grad_K;
    const int D = st->D, Dp = st->Dp;
    Dev dv{ctx, Dp, D};
    const size_t mb = (size_t)Dp * Dp * 16;
    const size_t mcount = (size_t)Dp * Dp * 2;
    const int nn = 2 * ip->n_quad;
    const double s2 = std::sqrt(2.0);
    const double coef[3] = {1.0, 1.0, r.g};
    // ----------------------------------------------------------------------------------
    // pullback (src/observables.jl:161-199) in reduced channels
    // ----------------------------------------------------------------------------------
    ck(hipEventRecord(ctx->ev[2], ctx->stream));
    ck(ctx->tmpM.ensure((size_t)nn * mb));
    ck(ctx->scratch.ensure(12 * mb));
    double* M = ctx->tmpM.d();
    double* cov = ctx->scratch.d();
    double* env = cov + mcount;
    double* T1 = cov + 2 * mcount;
    double* T2 = cov + 3 * mcount;
    double* U = cov + 4 * mcount;
    double* X1 = cov + 5 * mcount;
    double* X2 = cov + 6 * mcount;
    double* cterm = cov + 7 * mcount;
    double* tmp = cov + 8 * mcount;
    double* tmp2 = cov + 9 * mcount;
    // dE = 2 Re tr(XR dR) + 2 Re tr(XQ dQ) for any (dR, dQ); covector = XR - XQ R^+
    double* XR = cov + 10 * mcount;
    double* XQ = cov + 11 * mcount;
    ck(hipMemsetAsync(env, 0, mb, ctx->stream));
    ck(hipMemsetAsync(XR, 0, 2 * mb, ctx->stream));
    const long long node = (long long)Dp * Dp;   // complex elements per matrix
    for (int term = 0; term < 3; ++term) {
        const int sf = 2 * term, sa = sf + 1;
        const int lf = kLive[sf], la = kLive[sa];
        const double* Fs = tb.samples[sf].d();
        const double* As = tb.samples[sa].d();
        const long long fstride = lf * node, astride = la * node;
        // adjoint sample of node i is at index nn-1-i: base at nn-1, negative stride
        auto Aptr = [&](int ch) { return As + 2 * ((long long)(nn - 1) * astride + ch * node); };
        auto Fptr = [&](int ch) { return Fs + 2 * (ch * node); };
        ck(hipMemsetAsync(T1, 0, 3 * mb, ctx->stream));   // T1, T2, U
        if (term < 2) {
            // k = 1 (F1 = rho):  T1 += rho R^+ (sum w A1), T2 += rho (sum w A1)
            hipLaunchKernelGGL(rc_lincomb, dv.eb(), dim3(256), 0, ctx->stream, Aptr(0), -astride,
                               ctx->nodes_w.d(), nn, dv.n(), 0.0, X1);
            hipLaunchKernelGGL(rc_lincomb, dv.eb(), dim3(256), 0, ctx->stream, Aptr(1), -astride,
                               ctx->nodes_j.d(), nn, dv.n(), 0.0, X2);           // sum j A2
            hipLaunchKernelGGL(rc_lincomb, dv.eb(), dim3(256), 0, ctx->stream, Fptr(0), fstride,
                               ctx->nodes_j.d(), nn, dv.n(), 0.0, tmp2);         // sum j F2
            ck(hipGetLastError());
            dv.gemm(0, st->Rh.d(), 0, X1, tmp);
            dv.gemm(0, st->rho.d(), 0, tmp, T1);
            dv.gemm(0, st->rho.d(), 0, X1, T2);
            // k = 2: F2 (w R^+ A2), w F2 A2 ; k = 3: F2^+ (w R^+ A2^+), w F2^+ A2^+
            for (int k3 = 0; k3 < 2; ++k3) {
                GemmArgs gb{};
                gb.opA = 0; gb.A = st->Rh.d(); gb.strideA = 0;
                gb.opB = k3; gb.B = Aptr(1); gb.strideB = -astride;
                gb.C = M; gb.strideC = node;
                gb.alpha = ctx->nodes_w.d(); gb.alpha_stride = 1;
                dv.batched(gb, nn);
                GemmArgs gr{};
                gr.opA = k3; gr.A = Fptr(0); gr.strideA = fstride;
                gr.opB = 0; gr.B = M; gr.strideB = node;
                dv.reduce(gr, nn, T1, 1.0, 1.0);
                GemmArgs g2{};
                g2.opA = k3; g2.A = Fptr(0); g2.strideA = fstride;
                g2.opB = k3; g2.B = Aptr(1); g2.strideB = -astride;
                g2.alpha = ctx->nodes_w.d(); g2.alpha_stride = 1;
                dv.reduce(g2, nn, T2, 1.0, 1.0);
            }
            // u = sqrt2 (rho sum j A2 + (sum j F2)^+), the coefficient of the edge operator
            dv.gemm(0, st->rho.d(), 0, X2, U, s2);
            dv.herm(1, s2, tmp2, tmp);
            dv.axpby(cd(1, 0), U, cd(1, 0), tmp, U);
            if (term == 0) {
                dv.axpby(cd(1, 0), T1, cd(1, 0), U, T1);           // op = R
            } else {
                // op = C = QR - RQ: dC = dQ R + Q dR - dR Q - R dQ
                dv.gemm(0, U, 0, st->Q.d(), T1, 1.0, 1.0);
                dv.gemm(0, st->Q.d(), 0, U, T1, -1.0, 1.0);
                dv.gemm(0, st->R.d(), 0, U, T2, 1.0, 1.0);
                dv.gemm(0, U, 0, st->R.d(), T2, -1.0, 1.0);
            }
        } else {
            // phi^4: sum_k F_k (w R^+ A_k + j k A_{k+1}) - (sum_k w F_k A_k) R^+,  F_1 = rho, A_5 = I
            for (int k = 1; k <= 4; ++k) {
                GemmArgs gb{};
                gb.opA = 0; gb.A = st->Rh.d(); gb.strideA = 0;
                gb.opB = 0; gb.B = Aptr(k - 1); gb.strideB = -astride;
                gb.C = M; gb.strideC = node;
                gb.alpha = ctx->nodes_w.d(); gb.alpha_stride = 1;
                gb.dalpha = ctx->nodes_j.d(); gb.dscale = k;
                if (k < 4) { gb.addD = Aptr(k); gb.strideD = -astride; }
                else gb.add_identity = 4.0;
                dv.batched(gb, nn);
                GemmArgs gr{};
                gr.opA = 0;
                gr.A = k == 1 ? st->rho.d() : Fptr(k - 2);
                gr.strideA = k == 1 ? 0 : fstride;
                gr.opB = 0; gr.B = M; gr.strideB = node;
                dv.reduce(gr, nn, T1, 1.0, 1.0);
                GemmArgs g2{};
                g2.opA = 0;
                g2.A = k == 1 ? st->rho.d() : Fptr(k - 2);
                g2.strideA = k == 1 ? 0 : fstride;
                g2.opB = 0; g2.B = Aptr(k - 1); g2.strideB = -astride;
                g2.alpha = ctx->nodes_w.d(); g2.alpha_stride = 1;
                dv.reduce(g2, nn, T2, 1.0, 1.0);
            }
        }
        // environment: hermitised adjoint channel 1 at the endpoint
        const double* aend = As + 2 * ((long long)nn * astride);
        dv.herm(0, 1.0, aend, tmp);
        const double c = coef[term];
        dv.axpby(cd(1, 0), XR, cd(c, 0), T1, XR);
        dv.axpby(cd(1, 0), XQ, cd(c, 0), T2, XQ);
        dv.axpby(cd(1, 0), env, cd(c, 0), tmp, env);
    }
    auto covector = [&] {   // horizontal directions dQ = -R^+ W
        dv.gemm(0, XQ, 0, st->Rh.d(), cterm);
        dv.axpby(cd(1, 0), XR, cd(-1, 0), cterm, cov);
    };
    covector();
    ck(hipEventRecord(ctx->ev[3], ctx->stream));
    if (raw_cov) pad_out(ctx, D, Dp, cov, raw_cov);
    if (raw_env) pad_out(ctx, D, Dp, env, raw_env);

    // ----------------------------------------------------------------------------------
    // finish_gradient (src/observables.jl:201-212)
    // ----------------------------------------------------------------------------------
    if (grad || grad_R || grad_K) {
        if (!(flags & RCMPS_FLAG_NO_BOUNDARY)) {
            ck(ctx->tmpM.ensure(2 * mb));
            double* hb = X1;
            // aims two digits below the reference's 1e-11 and accepts at its bar
            boundary_solve(ctx, st, env, 1e-13, 1e-11, hb, S);
            // rho0 responds to dL(rho0) = dQ rho0 + rho0 dQ^+ + dR rho0 R^+ + R rho0 dR^+:
            // XR += rho R^+ H, XQ += rho H  (covector += rho R^+ H - rho H R^+)
            dv.gemm(0, st->rho.d(), 0, st->Rh.d(), tmp);
            dv.gemm(0, tmp, 0, hb, XR, 1.0, 1.0);
            dv.gemm(0, st->rho.d(), 0, hb, XQ, 1.0, 1.0);
            covector();
        }
    }
    if (grad_R || grad_K) {
        // Q = -iK - R^+R/2: dQ = -i dK - (dR^+ R + R^+ dR)/2, so
        // dE = 2 Re tr[(XR - herm(XQ) R^+) dR] + 2 Re tr[herm(-i XQ) dK]; returned as the
        // gradients G with dE = Re <G, d> = Re sum conj(G_ij) d_ij (K restricted to Hermitian)
        if (grad_R) {
            dv.herm(0, 1.0, XQ, tmp2);
            dv.gemm(0, tmp2, 0, st->Rh.d(), tmp);
            dv.axpby(cd(1, 0), XR, cd(-1, 0), tmp, tmp);
            dv.herm(1, 2.0, tmp, tmp2);
            pad_out(ctx, D, Dp, tmp2, grad_R);
        }
        if (grad_K) {
            dv.herm(1, 1.0, XQ, tmp);
            dv.axpby(cd(0, -1), XQ, cd(0, 1), tmp, tmp2);
            pad_out(ctx, D, Dp, tmp2, grad_K);
        }
    }
    if (grad) {
        // G = [(rho + reg I)^{-1} cov]^+ = cov^+ (rho + reg I)^{-1}
        rcmps_state* ms = const_cast<rcmps_state*>(st);
        if (ms->metric_reg_cached != metric_reg) {
            Mat a = st->rho_host;
            for (int i = 0; i < D; ++i) a[(size_t)i * D + i] += metric_reg;
            Mat inv;
            if (!rc::inverse(D, a, inv)) throw Fail{RCMPS_ERR_SOLVER};
            ck(ms->Minv.ensure(mb));
            upload(ctx, D, Dp, inv, ms->Minv.d());
            ms->metric_reg_cached = metric_reg;
        }
        dv.gemm(1, cov, 0, st->Minv.d(), tmp);
        pad_out(ctx, D, Dp, tmp, grad);
    }
    ck(hipEventRecord(ctx->ev[4], ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    float pms = 0, bms = 0;
    ck(hipEventElapsedTime(&pms, ctx->ev[2], ctx->ev[3]));
    ck(hipEventElapsedTime(&bms, ctx->ev[3], ctx->ev[4]));
    S->pullback_ms = pms;
    S->boundary_ms = bms;
}

bool valid_integrator(const rcmps_integrator* ip) {
    return ip && ip->n_quad >= 8 && ip->max_point > 0 && ip->atol > 0 && ip->rtol > 0 && ip->maxiters >= 1;
}

// A batch of evaluations: grouped by padded size, one persistent launch per group, then the
// pullback of each evaluation.  Every request gets its stats; the first failure is returned.
int fg_phi4_batch_impl(rcmps_context* ctx, std::vector<EvalReq>& reqs, const rcmps_integrator* ip,
                       double metric_reg, int flags) {
    if (!ctx || !valid_integrator(ip) || !(metric_reg >= 0)) return RCMPS_ERR_ARGUMENT;
    if (std::sqrt(ip->max_point) >= RC_SOURCE_TMAX) return RCMPS_ERR_UNSUPPORTED;
    for (auto& r : reqs) {
        if (!r.st || !r.energy || !r.S || !std::isfinite(r.g)) return RCMPS_ERR_ARGUMENT;
        if (r.st->Dp / 16 > 4) return RCMPS_ERR_UNSUPPORTED;
        std::memset(r.S, 0, sizeof(*r.S));
    }
    auto t_start = std::chrono::steady_clock::now();
    int rc = guard([&] {
        ck(hipSetDevice(ctx->device));
        ensure_quadrature(ctx, ip);
        for (int Dp = 16; Dp <= 64; Dp += 16) {
            std::vector<int> idx;
            for (size_t i = 0; i < reqs.size(); ++i)
                if (reqs[i].st->Dp == Dp) idx.push_back((int)i);
            if (idx.empty()) continue;
            run_class(ctx, Dp, (flags & RCMPS_FLAG_SPLIT) != 0, ip, idx, reqs);
            for (size_t j = 0; j < idx.size(); ++j) {
                EvalReq& r = reqs[idx[j]];
                if (r.S->status != RCMPS_OK || !r.adjoint()) continue;
                const int code = guard([&] { finish_state(ctx, r, *ctx->tb[j], ip, metric_reg, flags); });
                if (code != RCMPS_OK) r.S->status = code;
            }
        }
    });
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    for (auto& r : reqs) {
        r.S->total_ms = total;
        if (rc && r.S->status == RCMPS_OK) r.S->status = rc;
    }
    if (rc) return rc;
    for (auto& r : reqs)
        if (r.S->status != RCMPS_OK) return r.S->status;
    return RCMPS_OK;
}

int fg_phi4_impl(rcmps_context* ctx, const rcmps_state* st, double g, const rcmps_integrator* ip,
                 double metric_reg, int flags, double* energy, double* grad, double* raw_cov, double* raw_env,
                 double* grad_R, double* grad_K, rcmps_stats* stats) {
    if (!ctx || !st || !ip || !energy) return RCMPS_ERR_ARGUMENT;
    rcmps_stats local{};
    std::vector<EvalReq> reqs(1);
    reqs[0].st = st; reqs[0].g = g; reqs[0].energy = energy;
    reqs[0].grad = grad; reqs[0].raw_cov = raw_cov; reqs[0].raw_env = raw_env;
    reqs[0].grad_R = grad_R; reqs[0].grad_K = grad_K;
    reqs[0].S = stats ? stats : &local;
    return fg_phi4_batch_impl(ctx, reqs, ip, metric_reg, flags);
}

}  // namespace

int rcmps_fg_phi4(rcmps_context* ctx, const rcmps_state* st, double g, const rcmps_integrator* ip,
                  double metric_reg, int flags, double* energy, double* grad, double* raw_cov, double* raw_env,
                  rcmps_stats* stats) {
    return fg_phi4_impl(ctx, st, g, ip, metric_reg, flags, energy, grad, raw_cov, raw_env, nullptr, nullptr,
                        stats);
}

int rcmps_fg_phi4_batch(rcmps_context* ctx, int n, const rcmps_state* const* states, const double* g,
                        const rcmps_integrator* ip, double metric_reg, int flags, double* energies,
                        double* const* grads, rcmps_stats* stats) {
    if (!ctx || n < 1 || !states || !g || !energies) return RCMPS_ERR_ARGUMENT;
    std::vector<rcmps_stats> local(stats ? 0 : n);
    std::vector<EvalReq> reqs(n);
    for (int i = 0; i < n; ++i) {
        reqs[i].st = states[i];
        reqs[i].g = g[i];
        reqs[i].energy = energies + i;
        reqs[i].grad = grads ? grads[i] : nullptr;
        reqs[i].S = stats ? stats + i : &local[i];
    }
    return fg_phi4_batch_impl(ctx, reqs, ip, metric_reg, flags);
}

int rcmps_fg_phi4_rk(rcmps_context* ctx, const rcmps_state* st, double g, const rcmps_integrator* ip, int flags,
                     double* energy, double* grad_R, double* grad_K, rcmps_stats* stats) {
    return fg_phi4_impl(ctx, st, g, ip, 0.0, flags, energy, nullptr, nullptr, nullptr, grad_R, grad_K, stats);
}

int rcmps_trajectory_samples(rcmps_context* ctx, int system, double* dst, long long capacity) {
    if (!ctx || system < 0 || system >= RC_NSYS || !dst || ctx->last_Dp == 0 || ctx->tb.empty())
        return RCMPS_ERR_ARGUMENT;
    if (system % 2 && !ctx->last_adjoint) return RCMPS_ERR_ARGUMENT;
    const int D = ctx->last_D, Dp = ctx->last_Dp, nl = kLive[system];
    const long long need = (long long)ctx->last_npts * nl * D * D * 2;
    if (capacity < need) return RCMPS_ERR_ARGUMENT;
    return guard([&] {
        const double* src = ctx->tb[0]->samples[system].d();
        const int rows = ctx->last_npts * nl;
        for (int r = 0; r < rows; ++r)
            pad_out(ctx, D, Dp, src + (size_t)r * Dp * Dp * 2, dst + (size_t)r * D * D * 2);
        ck(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
