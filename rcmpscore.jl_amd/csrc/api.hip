// C-ABI of librcmps_hip.so (declared in include/rcmps.h): state preparation, fg_phi4 and the
// boundary Poisson solve on one MI355X.  All O(D^3)-per-iteration work runs on the GPU; the host
// holds the Krylov bookkeeping (Hessenberg columns, Givens rotations) and, once per state, the
// D x D spectral factorisation of Q that the no-jump preconditioner applies.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rcmps.h"
#include "excitation.h"
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
// Arnoldi basis of L or L* on one right-hand side (src/kinetic_krylov.jl:18-33)
struct KrylovBasis {
    DevBuf V;                 // (cap + 1) vectors of n = Dp^2 complex
    std::vector<cd> H;        // (cap + 1) x cap, row-major
    int n = 0, cap = 0, k = 0;
    double beta = 0.0;
    bool exhausted = false, started = false;
};

// workspace of the infinite-line kinetic backend
struct KineticWork {
    KrylovBasis basis[4];     // (op R, op [Q,R]) x (forward, adjoint)
    DevBuf X, A, X2, A2, rhs2, w4, hd4, LX, T, Yd, zd, wd, wzd, res, hd, tmp;
    int actions = 0;
    double max_res = 0.0;
    double ms[4] = {0.0, 0.0, 0.0, 0.0};   // Arnoldi, projected solves, solutions + residuals, total
    void release() {
        for (auto& b : basis) { b.V.release(); b = KrylovBasis(); }
        DevBuf* all[] = {&X, &A, &X2, &A2, &rhs2, &w4, &hd4, &LX, &T, &Yd, &zd, &wd, &wzd, &res, &hd, &tmp};
        for (auto* d : all) d->release();
    }
};

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

// square root of a computed squared norm: rounding can leave it slightly negative (clamped to 0), and
// a NaN stays NaN so that the callers' finiteness checks see it
inline double norm_sqrt(double v) { return std::isnan(v) ? v : std::sqrt(std::max(0.0, v)); }

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
// Gauss-Legendre nodes and weights on (-1, 1): Newton on P_n from the Chebyshev-like guess.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
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
}

void quadrature(int n, double max_point, std::vector<double>& pts, std::vector<double>& wts) {
    std::vector<double> x, w;
    gauss_legendre(n, x, w);
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

// device address of this device's copy of the no-alloc scratch array (defined at the end of this file)
hipError_t noalloc_scratch_setup(int device);

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
    // GMRES read-back: two pinned slots of one Hessenberg column each, and their completion events
    double* kry_pinned = nullptr;
    size_t kry_pinned_doubles = 0;
    hipEvent_t kry_ev[2] = {nullptr, nullptr};
    int last_D = 0, last_Dp = 0, last_npts = 0;
    int last_tb = 0;             // trajectory buffers of the last evaluation's first state
    bool last_adjoint = false;   // the last evaluation integrated the adjoint systems
    // quadrature cache: nodes (plus endpoint) and the pullback weights on the device
    int q_n = 0;
    double q_max_point = -1.0;
    std::vector<double> q_pts;
    hipEvent_t ev[8];
    double clock_mhz = 100.0;   // constant-rate device clock
    // infinite-line kinetic backend: controls (KineticOptions / KineticKrylovOptions defaults) and workspace
    rcmps_kinetic_options kopt{32, 512, 1e-11, 1e-10, 1e-12, 1e-12, 16, 512};
    KineticWork kin;
    // batched-GEMM tile variant per (padded size, batch class ceil(log2 nbatch)): measured on first
    // use unless forced (gemm_force >= 0)
    struct GemmTune {
        int variant = 0;
        double us[4] = {0.0, 0.0, 0.0, 0.0};
    };
    int gemm_force = -1;
    // set inside rcmps_boundary_environment_noalloc: products take the tuned variant or variant 0
    // (same bits) instead of measuring, since the tuner allocates its operands
    bool no_alloc = false;
    std::map<std::pair<int, int>, GemmTune> gemm_tuned;
    DevBuf tune_buf;
    // contexts (own stream and workspaces) on which the per-state pullback, boundary and metric
    // solves of a batch run concurrently, one host thread each
    std::vector<rcmps_context*> workers;
    // states created on this context whose eigenvector condition estimate of Q exceeds this apply
    // the no-jump inverse in the Schur basis instead of the eigenbasis
    double nojump_kappa = 1e6;
    // layout of the D x D matrices at the C ABI (RCMPS_LAYOUT_ROW_MAJOR or RCMPS_LAYOUT_COLUMN_MAJOR)
    int layout = RCMPS_LAYOUT_ROW_MAJOR;
    // rocBLAS handle on this context's stream, for the dense transfer resolvents (created on first use)
    rocblas_handle blas = nullptr;
};

struct rcmps_state {
    int D = 0, Dp = 0;
    // Z2 symmetry under diag(1, .., 1, -1, .., -1) (D/2 each): R odd, K and rho0 even, exactly.  When
    // the parity blocks are whole 16 x 16 tiles, z2_tiles = D / 32 (tiles per parity block), else 0.
    int z2_tiles = 0;
    DevBuf R, K, Q, Qh, Rh, C, Ch, rho, V, Vinv, lam, ident;
    DevBuf VH, VinvH;   // V^+ and V^-+ for the fused two-sided products of the eigenbasis form
    // eigendecomposition L = W diag(mu) W^-1 of the transfer superoperator on the padded Dp^2 space
    // (W, W^-1 column-major, identity on the padded indices), set by rcmps_state_set_transfer_eigen:
    // the excitation kinetic resolvents (1 - z (L - s))^-1 are then products with W and W^-1
    int tN = 0;
    int tzero = -1;      // padded index of the zero mode (the stationary state), -1 if not isolated
    DevBuf tW, tWinv, tmu;
    // no-jump inverse: eigenbasis (V = eigenvectors of Q, Vinv their inverse, lam the eigenvalues) or,
    // when nojump_schur, the Schur basis (V = Z, Vinv = Z^+, Tq = T upper triangular, Q = Z T Z^+)
    int nojump_schur = 0;
    double nojump_kappa = 0.0;   // condition estimate |W|_F |W^-1|_F of the triangular eigenvectors
    DevBuf Tq;
    double shift = 0.0;
    std::vector<cd> rho_host;
    double metric_reg_cached = -1.0;
    DevBuf Minv;   // (rho + reg I)^{-1}
};

namespace {

// ------------------------------------------------------------------------------------------
// batched-GEMM tile variants and their autotuner.  Variant 0: one 16x16 output tile per
// workgroup of four waves splitting the k-steps (LDS reduction); 1: four tiles per workgroup, one
// per wave; 2: one tile per single-wave workgroup; 3: a 32x32 block per workgroup with both
// operand panels staged in LDS.  All of them produce the same bits (the per-wave variants keep
// the four k-step classes apart and sum them in variant 0's order), so the choice is purely a
// speed decision: the first product of a (padded size, batch class) on a context times each
// variant on operands of that shape and keeps the fastest.
// ------------------------------------------------------------------------------------------
constexpr int kGemmVariants = 4;

size_t gemm_lds_bytes(int Dp) { return (size_t)2 * 32 * (Dp + 1) * 16; }

int gemm_class(int nbatch) {
    int c = 0;
    while ((1 << c) < nbatch && c < 30) ++c;
    return c;
}

void launch_zgemm_variant(hipStream_t stream, const GemmArgs& g, int nbatch, int v) {
    const int nt = (g.Dp / 16) * (g.Dp / 16);
    if (v == 1) hipLaunchKernelGGL(rc_zgemm_batched_w<4>, dim3(ceil_div(nt, 4), nbatch), dim3(256), 0, stream, g);
    else if (v == 2) hipLaunchKernelGGL(rc_zgemm_batched_w<1>, dim3(nt, nbatch), dim3(64), 0, stream, g);
    else if (v == 3) {
        const int nb2 = (g.Dp / 16 + 1) / 2;
        hipLaunchKernelGGL(rc_zgemm_batched_lds, dim3(nb2 * nb2, nbatch), dim3(256), gemm_lds_bytes(g.Dp), stream, g);
    } else hipLaunchKernelGGL(rc_zgemm_batched, dim3(nt, nbatch), dim3(256), 0, stream, g);
}

int gemm_variant(rcmps_context* ctx, int Dp, int nbatch) {
    if (ctx->gemm_force >= 0) return ctx->gemm_force;
    const auto key = std::make_pair(Dp, gemm_class(nbatch));
    auto it = ctx->gemm_tuned.find(key);
    if (it != ctx->gemm_tuned.end()) return it->second.variant;
    if (ctx->no_alloc) return 0;
    // a stream under graph capture takes the default without measuring
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return 0;
    const int nb_t = std::min(1 << key.second, 1024);
    const size_t mat = (size_t)Dp * Dp * 2;   // doubles per matrix
    ck(ctx->tune_buf.ensure(3 * (size_t)nb_t * mat * 8));
    double* buf = ctx->tune_buf.d();
    ck(hipMemsetAsync(buf, 0, 3 * (size_t)nb_t * mat * 8, ctx->stream));
    GemmArgs g{};
    g.Dp = Dp; g.D = Dp; g.nbatch = nb_t;
    g.A = buf; g.B = buf + nb_t * mat; g.C = buf + 2 * nb_t * mat;
    g.strideA = g.strideB = g.strideC = (long long)Dp * Dp;
    hipEvent_t e0, e1;
    ck(hipEventCreate(&e0));
    ck(hipEventCreate(&e1));
    // three interleaved rounds of ten launches per variant, the best round of each counts; another
    // variant replaces the default only when it is at least 10% faster (timer noise never flips it)
    rcmps_context::GemmTune t;
    constexpr int reps = 10, rounds = 3;
    for (int v = 0; v < kGemmVariants; ++v) {
        launch_zgemm_variant(ctx->stream, g, nb_t, v);
        t.us[v] = INFINITY;
    }
    for (int round = 0; round < rounds; ++round)
        for (int v = 0; v < kGemmVariants; ++v) {
            ck(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < reps; ++r) launch_zgemm_variant(ctx->stream, g, nb_t, v);
            ck(hipEventRecord(e1, ctx->stream));
            ck(hipEventSynchronize(e1));
            float ms = 0.0f;
            ck(hipEventElapsedTime(&ms, e0, e1));
            t.us[v] = std::min(t.us[v], 1e3 * ms / reps);
        }
    for (int v = 1; v < kGemmVariants; ++v)
        if (t.us[v] < 0.9 * t.us[0] && t.us[v] < t.us[t.variant]) t.variant = v;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ck(hipGetLastError());
    ctx->gemm_tuned[key] = t;
    return t.variant;
}

void launch_zgemm(rcmps_context* ctx, const GemmArgs& g, int nbatch) {
    launch_zgemm_variant(ctx->stream, g, nbatch, gemm_variant(ctx, g.Dp, nbatch));
    ck(hipGetLastError());
}

// ------------------------------------------------------------------------------------------
// device helpers on Dp x Dp matrices
// ------------------------------------------------------------------------------------------
struct Dev {
    rcmps_context* ctx;
    int Dp, D;
    double* sm = nullptr;   // 64 doubles of device scalars; the context's `small` buffer when null
    int n() const { return Dp * Dp; }
    double* small() const;
    dim3 eb() const { return dim3(ceil_div(n(), 256)); }

    void gemm(int opA, const double* A, int opB, const double* B, double* C, double alpha = 1.0,
              double beta = 0.0) {
        GemmArgs g{};
        g.Dp = Dp; g.D = D; g.nbatch = 1; g.opA = opA; g.opB = opB;
        g.A = A; g.B = B; g.C = C;
        g.beta = beta;
        g.alpha_scale = alpha;
        launch_zgemm(ctx, g, 1);
    }
    void batched(GemmArgs g, int nbatch) {
        g.Dp = Dp; g.D = D; g.nbatch = nbatch;
        launch_zgemm(ctx, g, nbatch);
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
        double* out = small() + 32;
        hipLaunchKernelGGL(rc_dots, dim3(1), dim3(256), 0, ctx->stream, x, 0LL, n(), x, out);
        ck(hipGetLastError());
        const double v = fetch2(out).real();
        return norm_sqrt(v);
    }
};

double* Dev::small() const { return sm ? sm : ctx->small.d(); }

// Every D x D matrix that crosses the C ABI goes through these two copies: the caller's layout
// (row-major by default, column-major after rcmps_set_layout) <-> the padded row-major working layout.
void pad_in(rcmps_context* ctx, int D, int Dp, const double* src, double* dst) {
    if (ctx->layout == RCMPS_LAYOUT_COLUMN_MAJOR) {
        hipLaunchKernelGGL(rc_colmajor_in, dim3((Dp * Dp + 255) / 256), dim3(256), 0, ctx->stream, D, Dp, src, dst);
        ck(hipGetLastError());
        return;
    }
    ck(hipMemsetAsync(dst, 0, (size_t)Dp * Dp * 16, ctx->stream));
    ck(hipMemcpy2DAsync(dst, (size_t)Dp * 16, src, (size_t)D * 16, (size_t)D * 16, D, hipMemcpyDeviceToDevice,
                        ctx->stream));
}

void pad_out(rcmps_context* ctx, int D, int Dp, const double* src, double* dst) {
    if (ctx->layout == RCMPS_LAYOUT_COLUMN_MAJOR) {
        hipLaunchKernelGGL(rc_colmajor_out, dim3((D * D + 255) / 256), dim3(256), 0, ctx->stream, D, Dp, src, dst);
        ck(hipGetLastError());
        return;
    }
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
// no-jump inverse:  forward  X = P(Y) solves (Q - s/2) X + X (Q - s/2)^+ = -Y,
// adjoint X = P*(Y) solves (Q - s/2)^+ X + X (Q - s/2) = -Y  (src/transfer.jl:50-68), with the
// complex shift s/2 -> s/2 + i sigma of the excitation resolvents.  Q = V diag(lam) V^-1 makes the
// middle step elementwise; for a Q whose eigenvectors are ill-conditioned the state keeps its Schur
// form Q = Z T Z^+ instead (V = Z, V^-1 = Z^+) and the middle step is the triangular Sylvester
// recurrence.  Used as a right preconditioner, so the Krylov solutions are exact either way.
// ------------------------------------------------------------------------------------------
void nojump_middle(Dev& dv, const rcmps_state* st, bool adjoint, double s_re, double s_im, const double* Y,
                   double* Z) {
    if (st->nojump_schur) {
        hipLaunchKernelGGL(rc_sylvester_tri, dim3(1), dim3(256), rc_sylvester_tri_lds(dv.D), dv.ctx->stream, dv.Dp,
                           dv.D, adjoint ? 1 : 0, s_re, s_im, st->Tq.d(), Y, Z);
    } else if (s_im == 0.0) {
        hipLaunchKernelGGL(rc_sylvester_diag, dv.eb(), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, adjoint ? 1 : 0,
                           s_re, st->lam.d(), Y, Z);
    } else {
        hipLaunchKernelGGL(rc_sylvester_diag_c, dv.eb(), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, adjoint ? 1 : 0,
                           s_re, s_im, st->lam.d(), Y, Z);
    }
    ck(hipGetLastError());
}

void nojump_apply_c(Dev& dv, const rcmps_state* st, const double* Y, double* out, bool adjoint, double sigma,
                    double* t1, double* t2) {
    if (!st->nojump_schur && dv.Dp <= 64) {
        // two fused launches: (V^-1 Y V^-+ / den) then V (.) V^+ (adjoint: (V^+ Y V / den), V^-+ (.) V^-1)
        const int nbt = dv.Dp / 16;
        const double *a1 = adjoint ? st->VH.d() : st->Vinv.d(), *b1 = adjoint ? st->V.d() : st->VinvH.d();
        const double *a2 = adjoint ? st->VinvH.d() : st->V.d(), *b2 = adjoint ? st->Vinv.d() : st->VH.d();
        hipLaunchKernelGGL(rc_two_sided, dim3(nbt * nbt), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, a1, Y, b1, t1, 1,
                           adjoint ? 1 : 0, (const double*)st->lam.d(), st->shift, adjoint ? -sigma : sigma);
        hipLaunchKernelGGL(rc_two_sided, dim3(nbt * nbt), dim3(256), 0, dv.ctx->stream, dv.Dp, dv.D, a2,
                           (const double*)t1, b2, out, 0, 0, (const double*)nullptr, 0.0, 0.0);
        ck(hipGetLastError());
        return;
    }
    if (!adjoint) {
        dv.gemm(0, st->Vinv.d(), 0, Y, t1);
        dv.gemm(0, t1, 1, st->Vinv.d(), t2);
        nojump_middle(dv, st, false, st->shift, sigma, t2, t1);
        dv.gemm(0, st->V.d(), 0, t1, t2);
        dv.gemm(0, t2, 1, st->V.d(), out);
    } else {
        dv.gemm(1, st->V.d(), 0, Y, t1);
        dv.gemm(0, t1, 0, st->V.d(), t2);
        nojump_middle(dv, st, true, st->shift, -sigma, t2, t1);
        dv.gemm(1, st->Vinv.d(), 0, t1, t2);
        dv.gemm(0, t2, 0, st->Vinv.d(), out);
    }
}

void nojump_apply(Dev& dv, const rcmps_state* st, const double* Y, double* out, bool adjoint, double* t1,
                  double* t2) {
    nojump_apply_c(dv, st, Y, out, adjoint, 0.0, t1, t2);
}

// L(X) = Q X + X Q^+ + R X R^+ ; L*(X) = Q^+ X + X Q + R^+ X R
void lindblad(Dev& dv, const rcmps_state* st, const double* X, double* out, bool adjoint, double* t1) {
    if (dv.Dp <= 64) {   // one fused launch (rc_lindblad); four products otherwise
        const int nbt = dv.Dp / 16;
        hipLaunchKernelGGL(rc_lindblad, dim3(nbt * nbt), dim3(256), 0, dv.ctx->stream, dv.Dp, adjoint ? 1 : 0,
                           st->Q.d(), st->Qh.d(), st->R.d(), st->Rh.d(), X, out);
        ck(hipGetLastError());
        return;
    }
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
// host.  Solves A x = b; returns true when the true residual is <= tol_abs.  The Arnoldi vectors
// never depend on the host's Givens rotations, so step k + 1 is queued before the host reads step
// k's Hessenberg column (two pinned slots): the host's rotations overlap the next step's products.
// A step queued past convergence only fills a basis vector the solution does not use, so the
// iterates, the results and the matvec count are those of the unpipelined iteration.
// Buffers of one GMRES solve supplied by the caller instead of the context's: V ((m + 1) vectors), w
// (one vector), h (gmres_col(m) doubles + 64 bytes) on the device, and two host read-back slots of
// gmres_col(m) doubles each.
struct GmresBufs {
    double *V, *w, *h, *host;
};
inline size_t gmres_col(int m) { return (size_t)4 * (m + 1) + 2; }   // h1, h2 (complex), <w,w>

bool gmres(Dev& dv, const std::function<void(const double*, double*)>& A, const double* b, double* x,
           double tol_abs, int m, int maxiter, int& matvecs, double& resid, const GmresBufs* gb = nullptr) {
    rcmps_context* ctx = dv.ctx;
    const int n = dv.n();
    const size_t col = gmres_col(m);   // doubles per read-back
    double *V, *w, *hdev, *slots;
    if (gb) {
        V = gb->V;
        w = gb->w;
        hdev = gb->h;
        slots = gb->host;
    } else {
        ck(ctx->kry_V.ensure((size_t)(m + 1) * n * 16));
        ck(ctx->kry_w.ensure((size_t)n * 16));
        ck(ctx->kry_h.ensure(col * 8 + 64));
        if (ctx->kry_pinned_doubles < 2 * col) {
            ck(hipStreamSynchronize(ctx->stream));
            if (ctx->kry_pinned) ck(hipHostFree(ctx->kry_pinned));
            ctx->kry_pinned = nullptr;
            ctx->kry_pinned_doubles = 0;
            ck(hipHostMalloc(reinterpret_cast<void**>(&ctx->kry_pinned), 2 * col * 8));
            ctx->kry_pinned_doubles = 2 * col;
        }
        V = ctx->kry_V.d();
        w = ctx->kry_w.d();
        hdev = ctx->kry_h.d();
        slots = ctx->kry_pinned;
    }
    ck(hipMemsetAsync(x, 0, (size_t)n * 16, ctx->stream));
    std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), sn(m), yk(m);
    std::vector<double> cs(m);
    matvecs = 0;
    bool converged = false;
    auto residual_into_w = [&]() {
        A(x, w);
        ++matvecs;
        dv.axpby(cd(1, 0), b, cd(-1, 0), w, w);
        return dv.norm(w);
    };
    // Arnoldi step k: w = A v_k, classical Gram-Schmidt twice on the device (h1, h2: k + 1 complex
    // each, then <w,w>), v_{k+1} = w / |w|, and the column copied to pinned slot k & 1
    auto queue_step = [&](int k) {
        A(V + (size_t)k * n * 2, w);
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
        ck(hipMemcpyAsync(slots + (size_t)(k & 1) * col, hdev, col * 8, hipMemcpyDeviceToHost, ctx->stream));
        ck(hipEventRecord(ctx->kry_ev[k & 1], ctx->stream));
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
        queue_step(0);
        for (; k < m; ++k) {
            if (k + 1 < m) queue_step(k + 1);
            ck(hipEventSynchronize(ctx->kry_ev[k & 1]));
            ++matvecs;
            const cd* hv = reinterpret_cast<const cd*>(slots + (size_t)(k & 1) * col);
            for (int i = 0; i <= k; ++i)
                H[(size_t)i * m + k] += hv[i] + hv[(m + 1) + i];
            double hn = norm_sqrt(hv[2 * (m + 1)].real());
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
    const bool inv_ok = rc::upper_inverse(D, W, Winv);
    // condition estimate of the eigenvectors (Z is unitary): |W|_F |W^-1|_F >= cond_2(W)
    st->nojump_kappa = inv_ok ? fro(W) * fro(Winv) : INFINITY;
    if (!std::isfinite(st->nojump_kappa)) st->nojump_kappa = INFINITY;
    st->nojump_schur = (st->nojump_kappa > ctx->nojump_kappa && D <= RC_SYLVESTER_TRI_MAX_D) ? 1 : 0;
    if (!st->nojump_schur && !inv_ok) throw Fail{RCMPS_ERR_SOLVER};
    Mat V, Vinv;
    if (st->nojump_schur) {
        V = Z;
        Vinv = rc::adjoint(D, Z);
        ck(st->Tq.ensure((size_t)Dp * Dp * 16));
        upload(ctx, D, Dp, T, st->Tq.d());
        ck(hipFuncSetAttribute((const void*)rc_sylvester_tri, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)rc_sylvester_tri_lds(RC_SYLVESTER_TRI_MAX_D)));
    } else {
        rc::matmul(D, Z, W, V);
        rc::matmul(D, Winv, rc::adjoint(D, Z), Vinv);
    }
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
    ck(st->VH.ensure((size_t)Dp * Dp * 16));
    ck(st->VinvH.ensure((size_t)Dp * Dp * 16));
    upload(ctx, D, Dp, rc::adjoint(D, V), st->VH.d());
    upload(ctx, D, Dp, rc::adjoint(D, Vinv), st->VinvH.d());
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

// Tiles per parity block when R is exactly odd and K, rho0 exactly even under the split of the basis
// into halves, and the halves are whole tiles (D a multiple of 32); 0 otherwise.
int z2_tiles_of(int D, const Mat& r, const Mat& k, const Mat& rho) {
    if (D % 32 != 0) return 0;
    const int d = D / 2;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            const bool same = (i < d) == (j < d);
            const size_t e = (size_t)i * D + j;
            if (same && r[e] != cd(0, 0)) return 0;
            if (!same && (k[e] != cd(0, 0) || rho[e] != cd(0, 0))) return 0;
        }
    return d / 16;
}

// k-tile mask of a tile product A(b, :) B(:, c) whose operands have parities pa, pb (1: odd) under
// parity blocks of `zt` tiles (0: no structure, every k-tile).  The product is nonzero only when the
// output tile's parity pa ^ pb matches its blocks, and then only the k-tiles of block blk(b) ^ pa count.
int z2_kmask(int NB, int zt, int b, int c, int pa, int pb) {
    if (zt == 0) return 15;
    const int bb = b / zt, bc = c / zt;
    if ((bb ^ bc) != (pa ^ pb)) return 0;
    int m = 0;
    for (int t = 0; t < NB; ++t)
        if (t / zt == (bb ^ pa)) m |= 1 << t;
    return m;
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
        case RCMPS_ERR_SCRATCH_LIMIT: return "scratch_limit";
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
        for (auto& e : ctx->kry_ev) ck(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ck(ctx->small.ensure(4096));
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            ctx->clock_mhz = khz / 1000.0;
        ck(ctx->abort_word.ensure(256));
        ck(noalloc_scratch_setup(device));
        // the LDS-staged product variant takes up to 66 KB of dynamic LDS at Dp = 64
        ck(hipFuncSetAttribute(reinterpret_cast<const void*>(rc_zgemm_batched_lds),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)gemm_lds_bytes(64)));
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
    for (auto* w : ctx->workers) rcmps_context_destroy(w);
    ctx->workers.clear();
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf* all[] = {&ctx->traj_params, &ctx->traj_tasks, &ctx->traj_sys, &ctx->traj_sched,
                     &ctx->traj_points, &ctx->abort_word, &ctx->nodes_w, &ctx->nodes_j,
                     &ctx->tmpM, &ctx->partial, &ctx->scratch, &ctx->small, &ctx->kry_V, &ctx->kry_w,
                     &ctx->kry_t1, &ctx->kry_t2, &ctx->kry_x, &ctx->kry_b, &ctx->kry_z, &ctx->kry_h,
                     &ctx->kry_tmp, &ctx->tune_buf};
    for (auto* b : all) b->release();
    for (auto& b : ctx->tb) b->release();
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    ctx->kin.release();
    for (auto& e : ctx->ev) (void)hipEventDestroy(e);
    for (auto& e : ctx->kry_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->kry_pinned) (void)hipHostFree(ctx->kry_pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void rcmps_state_destroy(rcmps_state* st) {
    if (!st) return;
    DevBuf* all[] = {&st->R, &st->K, &st->Q, &st->Qh, &st->Rh, &st->C, &st->Ch, &st->rho,
                     &st->V, &st->Vinv, &st->lam, &st->ident, &st->Minv, &st->Tq, &st->VH, &st->VinvH,
                     &st->tW, &st->tWinv, &st->tmu};
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
            // the reference iterates its CPTP Arnoldi to tol / 10 (src/transfer.jl:103-105)
            bool okg = gmres(dv, A, ctx->kry_b.d(), ctx->kry_x.d(), steady_tol / 10, m, 200, mv, res);
            pre(ctx->kry_x.d(), st->rho.d());
            if (stats) { stats->krylov_matvecs = mv; stats->krylov_residual = res; }
            // accepted at steady_tol when the tighter target stagnates (as the boundary solve)
            if (!okg && !(res <= steady_tol)) throw Fail{RCMPS_ERR_SOLVER};
            rho_h = download(ctx, D, Dp, st->rho.d());
        }
        if (!checked_density(D, qh, rh, rho_h, steady_tol)) throw Fail{RCMPS_ERR_SOLVER};
        upload(ctx, D, Dp, rho_h, st->rho.d());
        st->rho_host = rho_h;
        st->z2_tiles = z2_tiles_of(D, rh, kh, rho_h);
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
// Device buffers of one boundary solve supplied by the caller instead of the context's: five
// matrices t1, t2, z, b, x, two matrices at tmp, 64 doubles at small, and the GMRES buffers.
struct BoundaryBufs {
    double *t1, *t2, *z, *b, *x, *tmp, *small;
    GmresBufs g;
};

void boundary_solve(rcmps_context* ctx, const rcmps_state* st, const double* a, double tol, double accept_tol,
                    double* h, rcmps_stats* stats, const BoundaryBufs* bb = nullptr) {
    const int D = st->D, Dp = st->Dp;
    const size_t mb = (size_t)Dp * Dp * 16;
    Dev dv{ctx, Dp, D};
    double *t1, *t2, *z, *b, *x, *tmp;
    if (bb) {
        t1 = bb->t1; t2 = bb->t2; z = bb->z; b = bb->b; x = bb->x; tmp = bb->tmp;
        dv.sm = bb->small;
    } else {
        ck(ctx->kry_t1.ensure(mb));
        ck(ctx->kry_t2.ensure(mb));
        ck(ctx->kry_z.ensure(mb));
        ck(ctx->kry_b.ensure(mb));
        ck(ctx->kry_x.ensure(mb));
        ck(ctx->kry_tmp.ensure(2 * mb));
        t1 = ctx->kry_t1.d(); t2 = ctx->kry_t2.d(); z = ctx->kry_z.d(); b = ctx->kry_b.d(); x = ctx->kry_x.d();
        tmp = ctx->kry_tmp.d();
    }
    double* tmp2 = tmp + (size_t)Dp * Dp * 2;
    double* tr2 = dv.small();
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
    bool ok = gmres(dv, A, b, x, tol_abs, m, 200, mv, res, bb ? &bb->g : nullptr);
    pre(x, h);
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
void build_state(rcmps_context* ctx, TrajBufs& tb, const rcmps_state* st, bool adjoint, bool quartic_only, bool z2,
                 int npts, std::vector<TrajSys>& sys, std::vector<TrajTask>& tasks) {
    const int D = st->D, Dp = st->Dp, NB = Dp / 16, NT = NB * NB;
    // parity of every live channel's stage input for a Z2 state (rho0 and I even; R, R^+, [Q,R] odd;
    // the insertions and the fixed sources flip it once per operator factor)
    static const int kParity[RC_NSYS][RC_MAX_LIVE] = {{1}, {0, 1}, {1}, {0, 1}, {1, 0, 1}, {0, 1, 0, 1}};
    const int zt = z2 ? st->z2_tiles : 0;
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
        // the commutator adjoint system (s = 3) forms channel 1's E = X C in NT tasks of its own,
        // ranked after the channels' tasks, so channel 1 has one level-1 product like every other
        // channel and channel 0's insertion does not wait for a second one
        Sy.ntasks = (!adjoint && s % 2) || (quartic_only && s < 4) ? 0 : kLive[s] * NT + (s == 3 ? NT : 0);
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
                    {
                        const int px = kParity[s][ch];
                        const int m1 = adj ? z2_kmask(NB, zt, b, c, px, 1) : z2_kmask(NB, zt, b, c, 1, px);
                        const int m2a = z2_kmask(NB, zt, b, c, 0, px), m2b = z2_kmask(NB, zt, b, c, px, 0);
                        const int m3 = adj ? z2_kmask(NB, zt, b, c, 1, px ^ 1) : z2_kmask(NB, zt, b, c, px ^ 1, 1);
                        t.kmask = m1 | (m2a << 4) | (m2b << 8) | (m3 << 12);
                    }
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
        if (s == 3 && Sy.ntasks > 0)
            for (int b = 0; b < NB; ++b)
                for (int c = 0; c < NB; ++c) {
                    TrajTask t{};
                    t.sys = sys_base + s; t.ch = 1; t.b = b; t.c = c; t.adj = 1;
                    t.edge_src = -1;
                    t.has_E = 1;
                    t.e_only = 1;
                    t.kmask = z2_kmask(NB, zt, b, c, kParity[3][1], 1) << 16;   // E = X C of channel 1
                    tasks.push_back(t);
                }
        sys.push_back(Sy);
    }
}

// Integrate every system of the given evaluations (all of one padded size Dp) in one persistent
// launch; states are packed into rounds of at most the resident workgroup count.  Evaluation j
// keeps its trajectories in ctx->tb[tb0 + j].
void run_class(rcmps_context* ctx, int Dp, bool split, bool quartic_only, bool timers, bool z2,
               const rcmps_integrator* ip, const std::vector<int>& idx, size_t tb0, std::vector<EvalReq>& reqs) {
    const int NB = Dp / 16, npts = 2 * ip->n_quad + 1;
    std::vector<TrajSys> sys;
    std::vector<TrajTask> tasks;
    while (ctx->tb.size() < tb0 + idx.size()) ctx->tb.emplace_back(new TrajBufs());
    std::vector<int> state_first, state_ntasks;
    for (size_t j = 0; j < idx.size(); ++j) {
        const int first = (int)tasks.size();
        build_state(ctx, *ctx->tb[tb0 + j], reqs[idx[j]].st, reqs[idx[j]].adjoint(), quartic_only, z2, npts, sys,
                    tasks);
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
    P.timers = timers ? 1 : 0;
    P.a = std::sqrt(ip->max_point);
    P.atol = ip->atol;
    P.rtol = ip->rtol;
    P.h0 = ip->initial_step > 0 ? ip->initial_step : 0.01;
    P.max_steps = ip->maxiters;
    // a single wait longer than a minute means a partner is not running: abort with WAIT_LIMIT
    P.wait_ticks = (long long)(60.0 * ctx->clock_mhz * 1e6);
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
        ck(hipMemsetAsync(ctx->tb[tb0 + j]->flags.p, 0, ctx->tb[tb0 + j]->flag_bytes, ctx->stream));
        ck(hipMemsetAsync(ctx->tb[tb0 + j]->out.p, 0, RC_NSYS * RC_OUT_STRIDE * 8, ctx->stream));
    }
    ck(hipMemsetAsync(ctx->abort_word.p, 0, 256, ctx->stream));
    ck(hipEventRecord(ctx->ev[0], ctx->stream));
    ck(rc_launch_trajectory(NB, split ? 1 : 0, nblocks, static_cast<const TrajParams*>(ctx->traj_params.p),
                            ctx->stream));
    ck(hipEventRecord(ctx->ev[1], ctx->stream));
    std::vector<double> outs(idx.size() * RC_NSYS * RC_OUT_STRIDE);
    for (size_t j = 0; j < idx.size(); ++j)
        ck(hipMemcpyAsync(outs.data() + j * RC_NSYS * RC_OUT_STRIDE, ctx->tb[tb0 + j]->out.d(),
                          RC_NSYS * RC_OUT_STRIDE * 8, hipMemcpyDeviceToHost, ctx->stream));
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
        ctx->last_tb = (int)tb0;
        ctx->last_adjoint = reqs[idx[0]].adjoint();
    }
}

// (rho0 + reg I)^{-1} on the device, cached per state and regulariser (observables.jl:210-211)
const double* metric_inverse(rcmps_context* ctx, const rcmps_state* st, double metric_reg) {
    rcmps_state* ms = const_cast<rcmps_state*>(st);
    if (ms->metric_reg_cached != metric_reg) {
        const int D = st->D;
        Mat a = st->rho_host;
        for (int i = 0; i < D; ++i) a[(size_t)i * D + i] += metric_reg;
        Mat inv;
        if (!rc::inverse(D, a, inv)) throw Fail{RCMPS_ERR_SOLVER};
        ck(ms->Minv.ensure((size_t)st->Dp * st->Dp * 16));
        upload(ctx, D, st->Dp, inv, ms->Minv.d());
        ms->metric_reg_cached = metric_reg;
    }
    return st->Minv.d();
}

// Pullback over the Gauss nodes, boundary correction and gradients of one evaluation whose
// trajectories are in tb (src/observables.jl:161-212).  With kin_cov/kin_env (device) the two
// kinetic terms come from the infinite-line backend and only the :phi^4: trajectories are pulled back.
void finish_state(rcmps_context* ctx, EvalReq& r, TrajBufs& tb, const rcmps_integrator* ip, double metric_reg,
                  int flags, const double* kin_cov = nullptr, const double* kin_env = nullptr) {
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
    for (int term = kin_cov ? 2 : 0; term < 3; ++term) {
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
            hipLaunchKernelGGL(rc_lincomb, dim3(ceil_div(dv.n(), 64)), dim3(64), 0, ctx->stream, Aptr(0), -astride,
                               ctx->nodes_w.d(), nn, dv.n(), 0.0, X1);
            hipLaunchKernelGGL(rc_lincomb, dim3(ceil_div(dv.n(), 64)), dim3(64), 0, ctx->stream, Aptr(1), -astride,
                               ctx->nodes_j.d(), nn, dv.n(), 0.0, X2);           // sum j A2
            hipLaunchKernelGGL(rc_lincomb, dim3(ceil_div(dv.n(), 64)), dim3(64), 0, ctx->stream, Fptr(0), fstride,
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
    if (kin_env) dv.axpby(cd(1, 0), env, cd(1, 0), kin_env, env);
    auto covector = [&] {   // horizontal directions dQ = -R^+ W
        dv.gemm(0, XQ, 0, st->Rh.d(), cterm);
        dv.axpby(cd(1, 0), XR, cd(-1, 0), cterm, cov);
        if (kin_cov) dv.axpby(cd(1, 0), cov, cd(1, 0), kin_cov, cov);
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
        dv.gemm(1, cov, 0, metric_inverse(ctx, st, metric_reg), tmp);
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

// ------------------------------------------------------------------------------------------
// infinite-line kinetic terms: angular resolvent with matrix-free Arnoldi
// (src/kinetic.jl:19-124, src/kinetic_krylov.jl:1-189)
// ------------------------------------------------------------------------------------------

// Gauss-Legendre nodes on (0, pi/2) with weights summing to one (kinetic.jl:19-23)
void kinetic_grid(int n, std::vector<double>& theta, std::vector<double>& w) {
    std::vector<double> x;
    gauss_legendre(n, x, w);
    theta.resize(n);
    for (int i = 0; i < n; ++i) {
        theta[i] = M_PI / 4 * (x[i] + 1);
        w[i] /= 2;
    }
}

// out_b = L(X_b) (or L*(X_b)) for nb Dp x Dp matrices at stride Dp^2; T is nb matrices of scratch
void lindblad_batched(Dev& dv, const rcmps_state* st, const double* X, double* out, double* T, int nb, bool adj) {
    const long long node = (long long)dv.Dp * dv.Dp;
    GemmArgs a{};   // out = Q X (forward), Q^+ X (adjoint)
    a.A = adj ? st->Qh.d() : st->Q.d(); a.B = X; a.strideB = node; a.C = out; a.strideC = node;
    dv.batched(a, nb);
    GemmArgs b{};   // out += X Q^+ (forward), X Q (adjoint)
    b.A = X; b.strideA = node; b.B = adj ? st->Q.d() : st->Qh.d(); b.C = out; b.strideC = node; b.beta = 1.0;
    dv.batched(b, nb);
    GemmArgs c{};   // T = R X (forward), X R (adjoint)
    if (!adj) { c.A = st->R.d(); c.B = X; c.strideB = node; }
    else { c.A = X; c.strideA = node; c.B = st->R.d(); }
    c.C = T; c.strideC = node;
    dv.batched(c, nb);
    GemmArgs d{};   // out += T R^+ (forward), R^+ T (adjoint)
    if (!adj) { d.A = T; d.strideA = node; d.B = st->Rh.d(); }
    else { d.A = st->Rh.d(); d.B = T; d.strideB = node; }
    d.C = out; d.strideC = node; d.beta = 1.0;
    dv.batched(d, nb);
}

void basis_grow(KrylovBasis& b, int n, int cap, hipStream_t stream) {
    if (n != b.n) {   // a state of another size: the stored vectors are of no use
        b.V.release();
        b.H.clear();
        b.cap = 0;
        b.k = 0;
        b.n = n;
    }
    if (cap <= b.cap) return;
    DevBuf nv;
    ck(nv.ensure((size_t)(cap + 1) * n * 16));
    if (b.cap > 0 && b.k + 1 > 0)
        ck(hipMemcpyAsync(nv.p, b.V.p, (size_t)std::min(b.k + 1, b.cap + 1) * n * 16, hipMemcpyDeviceToDevice, stream));
    ck(hipStreamSynchronize(stream));
    b.V.release();
    b.V = nv;
    std::vector<cd> H((size_t)(cap + 1) * cap, cd(0, 0));
    for (int i = 0; i <= b.cap && b.cap > 0; ++i)
        for (int j = 0; j < b.cap; ++j) H[(size_t)i * cap + j] = b.H[(size_t)i * b.cap + j];
    b.H.swap(H);
    b.cap = cap;
}

// Start a basis on rhs (device, Dp^2 complex): V_0 = rhs / |rhs|
void basis_start(rcmps_context* ctx, Dev& dv, KrylovBasis& b, const double* rhs, int cap) {
    b.k = 0;
    b.exhausted = false;
    b.started = true;
    basis_grow(b, dv.n(), cap, ctx->stream);
    b.beta = dv.norm(rhs);
    if (!std::isfinite(b.beta)) throw Fail{RCMPS_ERR_SOLVER};
    if (b.beta == 0.0) { b.exhausted = true; return; }
    dv.axpby(cd(1.0 / b.beta, 0), rhs, cd(0, 0), nullptr, b.V.d());
}

// Arnoldi steps of up to four bases at once, basis i up to targets[i] vectors: classical Gram-Schmidt
// twice, H from both passes (kinetic_krylov.jl:68-92); exhaustion at a vanishing new direction or the
// full dimension D^2.  Every launch serves all bases still extending (rc_lindblad_multi, rc_arn_*), and
// each basis gets the vectors, Hessenberg columns and exhaustion it gets when extended alone.  No step
// waits for the host: each writes its Hessenberg column (|w|^2 before, the two passes' projections,
// |w|^2 after) to its own slot and normalises on the device, and the host reads all columns once.  A
// step past an exhaustion divides by a vanishing norm; its vector and column are discarded (the basis
// keeps the vectors up to the exhaustion), so the kept bits are those of a step-by-step loop.
void basis_extend_multi(rcmps_context* ctx, const rcmps_state* st, Dev& dv, KrylovBasis* const* bs, const bool* adj,
                        const int* targets, int nb) {
    KineticWork& kw = ctx->kin;
    const int n = dv.n();
    const int D2 = st->D * st->D;
    int tg[4], k0[4], nsteps[4], maxsteps = 0;
    size_t stride[4], off[4], total = 0;
    for (int i = 0; i < nb; ++i) {
        tg[i] = std::min(targets[i], D2);
        basis_grow(*bs[i], n, tg[i], ctx->stream);
        k0[i] = bs[i]->k;
        nsteps[i] = std::max(0, tg[i] - k0[i]);
        stride[i] = 2 * (size_t)(bs[i]->cap + 1) + 2;   // complex entries per step's column slot
        off[i] = total;
        total += (size_t)nsteps[i] * stride[i];
        maxsteps = std::max(maxsteps, nsteps[i]);
    }
    if (maxsteps == 0) return;
    ck(kw.hd4.ensure(total * 16));
    ck(kw.w4.ensure((size_t)nb * n * 16));
    for (int s = 0; s < maxsteps; ++s) {
        RcArnBatch B{};
        RcLindbladMulti L{};
        B.n = n;
        int maxj = 0;
        for (int i = 0; i < nb; ++i) {
            if (s >= nsteps[i]) continue;
            const int j = k0[i] + s;
            RcArnBasis& a = B.b[B.nb];
            a.V = bs[i]->V.d();
            a.w = kw.w4.d() + 2 * (size_t)i * n;
            a.wl = a.w;
            a.h0 = kw.hd4.d() + 2 * (off[i] + (size_t)s * stride[i]);
            a.h1 = a.h0 + 2;
            a.h2 = a.h1 + 2 * (bs[i]->cap + 1);
            a.hn = a.h2 + 2 * (bs[i]->cap + 1);
            a.j = j;
            L.X[B.nb] = a.V + (size_t)j * n * 2;
            L.out[B.nb] = a.w;
            L.adj[B.nb] = adj[i] ? 1 : 0;
            maxj = std::max(maxj, j);
            ++B.nb;
        }
        if (dv.Dp <= 64) {
            const int nbt = dv.Dp / 16;
            hipLaunchKernelGGL(rc_lindblad_multi, dim3(nbt * nbt, B.nb), dim3(256), 0, ctx->stream, dv.Dp, st->Q.d(),
                               st->Qh.d(), st->R.d(), st->Rh.d(), L);
        } else {
            for (int q = 0; q < B.nb; ++q) lindblad(dv, st, L.X[q], L.out[q], L.adj[q] != 0, ctx->kry_t1.d());
        }
        hipLaunchKernelGGL(rc_arn_dots, dim3(maxj + 2, B.nb), dim3(256), 0, ctx->stream, B, 1);
        hipLaunchKernelGGL(rc_arn_subtract, dim3(ceil_div(n, 256), B.nb), dim3(256), 0, ctx->stream, B, 1);
        hipLaunchKernelGGL(rc_arn_dots, dim3(maxj + 1, B.nb), dim3(256), 0, ctx->stream, B, 2);
        hipLaunchKernelGGL(rc_arn_subtract, dim3(ceil_div(n, 256), B.nb), dim3(256), 0, ctx->stream, B, 2);
        hipLaunchKernelGGL(rc_arn_normalize, dim3(1, B.nb), dim3(256), 0, ctx->stream, B);
        ck(hipGetLastError());
    }
    std::vector<cd> hv(total);
    ck(hipMemcpyAsync(hv.data(), kw.hd4.d(), total * 16, hipMemcpyDeviceToHost, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nb; ++i) {
        KrylovBasis& b = *bs[i];
        for (int j = k0[i]; j < tg[i]; ++j) {
            const cd* col = hv.data() + off[i] + (size_t)(j - k0[i]) * stride[i];
            ++kw.actions;
            const double oldnorm = norm_sqrt(col[0].real());
            for (int r = 0; r <= j; ++r) b.H[(size_t)r * b.cap + j] += col[1 + r] + col[1 + (b.cap + 1) + r];
            const double beta = norm_sqrt(col[1 + 2 * (b.cap + 1)].real());
            if (!std::isfinite(oldnorm) || !std::isfinite(beta)) throw Fail{RCMPS_ERR_SOLVER};
            b.H[(size_t)(j + 1) * b.cap + j] = beta;
            b.k = j + 1;
            if (beta <= 10 * 2.220446049250313e-16 * std::max(oldnorm, 2.220446049250313e-16) || j + 1 == D2) {
                b.exhausted = true;
                break;
            }
        }
    }
}

// Projected shifted solves (I - z H_k) y = beta e1 for every node, by Givens QR of the Hessenberg
// matrix; the Arnoldi residual estimate |z h_{k+1,k} y_k| / beta (kinetic_krylov.jl:95-113).
// Nodes i0 <= i < i1 of projected_solutions; false on a singular or nonfinite projected solve.
bool projected_nodes(const KrylovBasis& b, const std::vector<double>& z, int i0, int i1, std::vector<cd>& Y,
                     double& est) {
    const int k = b.k, cap = b.cap;
    est = 0.0;
    std::vector<cd> A((size_t)k * k), g(k);
    for (int i = i0; i < i1; ++i) {
        for (int r = 0; r < k; ++r)
            for (int c = 0; c < k; ++c)
                A[(size_t)r * k + c] = (r == c ? cd(1, 0) : cd(0, 0)) - z[i] * (c >= r - 1 ? b.H[(size_t)r * cap + c] : cd(0, 0));
        std::fill(g.begin(), g.end(), cd(0, 0));
        g[0] = b.beta;
        for (int c = 0; c + 1 < k; ++c) {
            const cd x = A[(size_t)c * k + c], y = A[(size_t)(c + 1) * k + c];
            const double ax = std::abs(x), ay = std::abs(y);
            if (ay == 0) continue;
            double cs;
            cd sn;
            if (ax == 0) { cs = 0; sn = std::conj(y) / ay; }
            else { const double h = std::hypot(ax, ay); cs = ax / h; sn = (x / ax) * std::conj(y) / h; }
            for (int col = c; col < k; ++col) {
                const cd p = A[(size_t)c * k + col], q = A[(size_t)(c + 1) * k + col];
                A[(size_t)c * k + col] = cs * p + sn * q;
                A[(size_t)(c + 1) * k + col] = -std::conj(sn) * p + cs * q;
            }
            const cd p = g[c], q = g[c + 1];
            g[c] = cs * p + sn * q;
            g[c + 1] = -std::conj(sn) * p + cs * q;
        }
        for (int r = k - 1; r >= 0; --r) {
            cd acc = g[r];
            for (int c = r + 1; c < k; ++c) acc -= A[(size_t)r * k + c] * Y[(size_t)c + (size_t)k * i];
            const cd d = A[(size_t)r * k + r];
            if (std::abs(d) == 0) return false;
            Y[(size_t)r + (size_t)k * i] = acc / d;
        }
        for (int r = 0; r < k; ++r)
            if (!std::isfinite(Y[(size_t)r + (size_t)k * i].real()) || !std::isfinite(Y[(size_t)r + (size_t)k * i].imag()))
                return false;
        est = std::max(est, std::abs(z[i] * b.H[(size_t)k * cap + (k - 1)] * Y[(size_t)(k - 1) + (size_t)k * i]) / b.beta);
    }
    return true;
}

// Projected shifted solves (I - z_i H_k) y_i = beta e_1 of every node by Givens QR, and the largest
// Arnoldi residual estimate |z_i h_{k+1,k} y_i[k-1]| / beta.  With threads > 1 the nodes are split
// over host threads (each node is independent; the maximum does not depend on the split).
bool projected_solutions(const KrylovBasis& b, const std::vector<double>& z, std::vector<cd>& Y, double& est,
                         int threads) {
    const int k = b.k, nz = (int)z.size();
    Y.assign((size_t)k * nz, cd(0, 0));
    est = 0.0;
    const int nt = std::max(1, std::min(threads, nz / 8));
    if (nt == 1) return projected_nodes(b, z, 0, nz, Y, est);
    std::vector<double> e(nt, 0.0);
    std::vector<char> ok(nt, 1);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            const int i0 = (int)((long long)nz * t / nt), i1 = (int)((long long)nz * (t + 1) / nt);
            ok[t] = projected_nodes(b, z, i0, i1, Y, e[t]);
        });
    for (auto& th : pool) th.join();
    for (int t = 0; t < nt; ++t) {
        if (!ok[t]) return false;
        est = std::max(est, e[t]);
    }
    return true;
}

// Shifted solutions X_i = (I - z_i L)^{-1} rhs (or L*) for all nodes, each accepted on its true
// relative residual (kinetic_krylov.jl:114-155).  X: nz matrices at stride Dp^2.  Every job (one
// basis, its right-hand side and direction) runs its own sequence of tests; the extensions the
// jobs ask for in one round run together (basis_extend_multi), so each job's bits are those of
// running it alone.
constexpr int kProjectedThreads = 16;

struct KinJob {
    KrylovBasis* b;
    const double* rhs;
    bool adj;
    double* X;
    int want;                // the basis size this job waits for
    bool done;
    bool form = false;       // the projected solves passed: form and check the node solutions
    bool terminal = false;
    std::vector<cd> Y;       // projected solutions, k per node
};

// The host part of one pass of a job's tests (kinetic_krylov.jl:114-155): the projected solves.  It
// leaves the job asking for a larger basis, or with projected solutions to form (form), or failed
// (neither: a terminal basis whose projected solves did not pass).
void kinetic_project(const rcmps_kinetic_options& o, int maxdim, int initdim, KinJob& jb,
                     const std::vector<double>& z, const std::vector<double>& zone) {
    KrylovBasis& b = *jb.b;
    jb.want = 0;
    jb.form = false;
    if (b.k == 0) {
        jb.want = initdim;
        return;
    }
    const int k = b.k;
    jb.terminal = b.exhausted || k >= maxdim;
    double est = 0;
    // the node of largest z alone first: when its projected residual (one term of the full test)
    // misses the tolerance on a basis that can grow, the basis extends without the other nodes' solves
    if (!jb.terminal && z.size() > 1) {
        if (!projected_solutions(b, zone, jb.Y, est, 1) || est > o.tolerance) {
            jb.want = std::min(2 * k, maxdim);
            return;
        }
    }
    const bool ok = projected_solutions(b, z, jb.Y, est, kProjectedThreads);
    if (ok && (est <= o.tolerance || jb.terminal)) {
        jb.form = true;
        return;
    }
    if (!jb.terminal) jb.want = std::min(2 * k, maxdim);
}

// The device part: form the node solutions of a job whose projected solves passed and accept them on
// their true relative residuals; otherwise the job asks for a larger basis (or fails when its basis
// cannot grow).
void kinetic_form(rcmps_context* ctx, const rcmps_state* st, Dev& dv, KinJob& jb, int nz, int maxdim) {
    KineticWork& kw = ctx->kin;
    const rcmps_kinetic_options& o = ctx->kopt;
    KrylovBasis& b = *jb.b;
    const int n = dv.n(), k = b.k;
    ck(kw.Yd.ensure(jb.Y.size() * 16));
    ck(hipMemcpyAsync(kw.Yd.p, jb.Y.data(), jb.Y.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(rc_basis_combine, dim3(ceil_div(n, 256), ceil_div(nz, 8)), dim3(256), 0, ctx->stream, b.V.d(),
                       (long long)n, k, kw.Yd.d(), nz, n, jb.X);
    ck(hipGetLastError());
    lindblad_batched(dv, st, jb.X, kw.LX.d(), kw.T.d(), nz, jb.adj);
    hipLaunchKernelGGL(rc_shifted_residual, dim3(nz), dim3(256), 0, ctx->stream, jb.X, kw.LX.d(), jb.rhs, kw.zd.d(), n,
                       kw.res.d());
    ck(hipGetLastError());
    std::vector<double> res(nz);
    ck(hipMemcpyAsync(res.data(), kw.res.d(), (size_t)nz * 8, hipMemcpyDeviceToHost, ctx->stream));
    ck(hipStreamSynchronize(ctx->stream));
    double worst = 0;
    for (double r2 : res) worst = std::max(worst, std::sqrt(r2) / b.beta);
    const bool fallback = jb.terminal && worst > o.tolerance && worst <= o.fallback_tolerance;
    if (std::isfinite(worst) && (worst <= o.tolerance || fallback)) {
        kw.max_res = std::max(kw.max_res, worst);
        jb.done = true;
        return;
    }
    if (b.exhausted || jb.terminal) throw Fail{RCMPS_ERR_SOLVER};
    jb.want = std::min(2 * k, maxdim);
}

// Every round: the projected solves of all pending jobs at once (one host thread per job), then their
// node solutions on the device one job after the other, then the extensions they ask for together.
void kinetic_solutions(rcmps_context* ctx, const rcmps_state* st, Dev& dv, std::vector<KinJob>& jobs,
                       const std::vector<double>& z) {
    KineticWork& kw = ctx->kin;
    const rcmps_kinetic_options& o = ctx->kopt;
    const int n = dv.n(), nz = (int)z.size();
    const int maxdim = std::min(o.maxdim, st->D * st->D), initdim = std::min(o.initialdim, maxdim);
    if (jobs.size() > 4) throw Fail{RCMPS_ERR_ARGUMENT};
    for (KinJob& jb : jobs) {
        jb.want = 0;
        jb.done = false;
        KrylovBasis& b = *jb.b;
        if (!b.started) basis_start(ctx, dv, b, jb.rhs, initdim);
        if (b.beta == 0.0) {
            ck(hipMemsetAsync(jb.X, 0, (size_t)nz * n * 16, ctx->stream));
            jb.done = true;
        }
    }
    ck(kw.zd.ensure((size_t)nz * 8));
    ck(kw.res.ensure((size_t)nz * 8));
    ck(hipMemcpyAsync(kw.zd.d(), z.data(), (size_t)nz * 8, hipMemcpyHostToDevice, ctx->stream));
    int zhard = 0;
    for (int i = 1; i < nz; ++i)
        if (std::fabs(z[i]) > std::fabs(z[zhard])) zhard = i;
    const std::vector<double> zone(1, z[zhard]);
    using clock = std::chrono::steady_clock;
    auto since = [](clock::time_point t0) { return std::chrono::duration<double, std::milli>(clock::now() - t0).count(); };
    for (;;) {
        std::vector<KinJob*> pending;
        for (KinJob& jb : jobs)
            if (!jb.done) pending.push_back(&jb);
        if (pending.empty()) return;
        auto t0 = clock::now();
        if (pending.size() == 1) {
            kinetic_project(o, maxdim, initdim, *pending[0], z, zone);
        } else {
            std::vector<std::thread> threads;
            std::vector<std::exception_ptr> errs(pending.size());
            for (size_t q = 0; q < pending.size(); ++q)
                threads.emplace_back([&, q] {
                    try {
                        kinetic_project(o, maxdim, initdim, *pending[q], z, zone);
                    } catch (...) {   // nothing may leave a host thread
                        errs[q] = std::current_exception();
                    }
                });
            for (auto& t : threads) t.join();
            for (auto& e : errs)
                if (e) std::rethrow_exception(e);
        }
        kw.ms[1] += since(t0);
        KrylovBasis* bs[4];
        bool adj[4];
        int tg[4], nw = 0;
        for (KinJob* jb : pending) {
            if (jb->form) {
                t0 = clock::now();
                kinetic_form(ctx, st, dv, *jb, nz, maxdim);
                kw.ms[2] += since(t0);
            } else if (jb->want == 0) {
                throw Fail{RCMPS_ERR_SOLVER};   // a terminal basis whose projected solves failed
            }
            if (jb->done) continue;
            bs[nw] = jb->b;
            adj[nw] = jb->adj;
            tg[nw] = jb->want;
            ++nw;
        }
        if (nw == 0) continue;
        t0 = clock::now();
        basis_extend_multi(ctx, st, dv, bs, adj, tg, nw);
        kw.ms[0] += since(t0);
    }
}

// One kinetic spec on the infinite line: value, covector and environment (kinetic_krylov.jl:156-189)
// from its node solutions, X the forward solves on rho op^+ and A the adjoint solves on op^+.
// Scratch: 12 Dp x Dp matrices at kw.tmp.
void kinetic_components(rcmps_context* ctx, const rcmps_state* st, Dev& dv, int commutator,
                        const std::vector<double>& theta, const std::vector<double>& w, bool gradient, double& value,
                        double* cov, double* env, const double* X, const double* A) {
    KineticWork& kw = ctx->kin;
    const int nz = (int)theta.size();
    const long long node = (long long)dv.n();
    const size_t mcount = (size_t)dv.n() * 2;
    std::vector<double> z(nz), wz(nz);
    for (int i = 0; i < nz; ++i) { z[i] = std::cos(theta[i]); wz[i] = w[i] * z[i]; }
    ck(kw.wd.ensure((size_t)nz * 8));
    ck(kw.wzd.ensure((size_t)nz * 8));
    ck(hipMemcpyAsync(kw.wd.d(), w.data(), (size_t)nz * 8, hipMemcpyHostToDevice, ctx->stream));
    ck(hipMemcpyAsync(kw.wzd.d(), wz.data(), (size_t)nz * 8, hipMemcpyHostToDevice, ctx->stream));
    double* t = kw.tmp.d();
    double* u = t + mcount;
    double* S1 = t + 2 * mcount;
    double* S2 = t + 3 * mcount;
    double* S3 = t + 4 * mcount;
    double* S4 = t + 5 * mcount;
    double* Sa = t + 6 * mcount;
    double* a = t + 7 * mcount;
    double* m = t + 8 * mcount;
    double* nn = t + 9 * mcount;
    double* v = t + 10 * mcount;
    double* x2 = t + 11 * mcount;
    const double* op = commutator ? st->C.d() : st->R.d();
    const double* opd = commutator ? st->Ch.d() : st->Rh.d();
    hipLaunchKernelGGL(rc_lincomb, dim3(ceil_div(dv.n(), 64)), dim3(64), 0, ctx->stream, X, node, kw.wd.d(), nz, dv.n(), 0.0, u);
    ck(hipGetLastError());
    dv.trace_prod(op, u, ctx->small.d());
    value = dv.fetch2(ctx->small.d()).real();   // sum_i w_i Re tr(op x_i)
    if (!std::isfinite(value)) throw Fail{RCMPS_ERR_NONFINITE};
    if (!gradient) return;
    // a_i = A_i^+
    GemmArgs mb{};   // T_i = R^+ A_i^+
    mb.A = st->Rh.d(); mb.opB = 1; mb.B = A; mb.strideB = node; mb.C = kw.T.d(); mb.strideC = node;
    dv.batched(mb, nz);
    GemmArgs s1{};   // S1 = sum wz x_i T_i
    s1.A = X; s1.strideA = node; s1.B = kw.T.d(); s1.strideB = node; s1.alpha = kw.wzd.d(); s1.alpha_stride = 1;
    dv.reduce(s1, nz, S1, 1.0, 0.0);
    GemmArgs s2{};   // S2 = sum wz x_i a_i
    s2.A = X; s2.strideA = node; s2.opB = 1; s2.B = A; s2.strideB = node; s2.alpha = kw.wzd.d(); s2.alpha_stride = 1;
    dv.reduce(s2, nz, S2, 1.0, 0.0);
    GemmArgs nb{};   // T_i = R x_i
    nb.A = st->R.d(); nb.B = X; nb.strideB = node; nb.C = kw.T.d(); nb.strideC = node;
    dv.batched(nb, nz);
    GemmArgs s3{};   // S3 = sum wz a_i T_i
    s3.opA = 1; s3.A = A; s3.strideA = node; s3.B = kw.T.d(); s3.strideB = node; s3.alpha = kw.wzd.d(); s3.alpha_stride = 1;
    dv.reduce(s3, nz, S3, 1.0, 0.0);
    GemmArgs s4{};   // S4 = sum wz a_i x_i
    s4.opA = 1; s4.A = A; s4.strideA = node; s4.B = X; s4.strideB = node; s4.alpha = kw.wzd.d(); s4.alpha_stride = 1;
    dv.reduce(s4, nz, S4, 1.0, 0.0);
    hipLaunchKernelGGL(rc_lincomb, dim3(ceil_div(dv.n(), 64)), dim3(64), 0, ctx->stream, A, node, kw.wd.d(), nz, dv.n(), 0.0, Sa);
    ck(hipGetLastError());
    dv.herm(1, 1.0, Sa, a);                                  // a = sum w a_i
    dv.gemm(0, S2, 0, st->Rh.d(), x2);                       // m = S1 - S2 R^+
    dv.axpby(cd(1, 0), S1, cd(-1, 0), x2, m);
    dv.gemm(0, st->R.d(), 0, S4, x2);                        // nn = S3 - R S4
    dv.axpby(cd(1, 0), S3, cd(-1, 0), x2, nn);
    dv.gemm(0, a, 0, st->rho.d(), v);                        // v = a rho
    dv.gemm(0, opd, 0, a, x2);                               // env = herm(op^+ a)
    dv.herm(0, 1.0, x2, env);
    if (commutator) {
        dv.gemm(0, u, 0, st->Q.d(), m, 1.0, 1.0);            // m += u Q - Q u - R u R^+ + u R R^+
        dv.gemm(0, st->Q.d(), 0, u, m, -1.0, 1.0);
        dv.gemm(0, st->R.d(), 0, u, x2);
        dv.gemm(0, x2, 0, st->Rh.d(), m, -1.0, 1.0);
        dv.gemm(0, u, 0, st->R.d(), x2);
        dv.gemm(0, x2, 0, st->Rh.d(), m, 1.0, 1.0);
        dv.gemm(0, st->Qh.d(), 0, v, nn, 1.0, 1.0);          // nn += Q^+ v - v Q^+ - R v R^+ + R R^+ v
        dv.gemm(0, v, 0, st->Qh.d(), nn, -1.0, 1.0);
        dv.gemm(0, st->R.d(), 0, v, x2);
        dv.gemm(0, x2, 0, st->Rh.d(), nn, -1.0, 1.0);
        dv.gemm(0, st->Rh.d(), 0, v, x2);
        dv.gemm(0, st->R.d(), 0, x2, nn, 1.0, 1.0);
    } else {
        dv.axpby(cd(1, 0), m, cd(1, 0), u, m);
        dv.axpby(cd(1, 0), nn, cd(1, 0), v, nn);
    }
    dv.herm(1, 1.0, nn, x2);                                 // cov = (m + nn^+) / 2
    dv.axpby(cd(0.5, 0), m, cd(0.5, 0), x2, cov);
}

// Finished Riemannian gradient of a covector/environment pair (observables.jl:201-212) into G (Dp x Dp)
void finish_gradient_dev(rcmps_context* ctx, const rcmps_state* st, Dev& dv, const double* cov, const double* env,
                         double metric_reg, double* G, double* scratch3, rcmps_stats* S) {
    const size_t mcount = (size_t)dv.n() * 2;
    double* h = scratch3;
    double* c2 = scratch3 + mcount;
    double* x = scratch3 + 2 * mcount;
    boundary_solve(ctx, st, env, 1e-13, 1e-11, h, S);
    dv.copy(cov, c2);
    dv.gemm(0, st->rho.d(), 0, st->Rh.d(), x);
    dv.gemm(0, x, 0, h, c2, 1.0, 1.0);
    dv.gemm(0, st->rho.d(), 0, h, x);
    dv.gemm(0, x, 0, st->Rh.d(), c2, -1.0, 1.0);
    dv.gemm(1, c2, 0, metric_inverse(ctx, st, metric_reg), G);
}

// kinetic_resolvent (kinetic.jl:77-124) over the selected kinetic terms (bit 0: kinetic1, bit 1:
// kinetic2) with weight one each: node doubling until energy and finished gradient settle.
// cov, env (device, Dp-padded) receive the accepted level's raw covector and environment.
void kinetic_resolvent_dev(rcmps_context* ctx, const rcmps_state* st, int which, double metric_reg, bool gradient,
                           double& value, double* cov, double* env, rcmps_stats* S) {
    const rcmps_kinetic_options& o = ctx->kopt;
    KineticWork& kw = ctx->kin;
    const int Dp = st->Dp, D = st->D;
    Dev dv{ctx, Dp, D};
    const size_t mb = (size_t)Dp * Dp * 16;
    const size_t mcount = (size_t)Dp * Dp * 2;
    for (auto& b : kw.basis) { b.k = 0; b.exhausted = false; b.started = false; b.beta = 0; std::fill(b.H.begin(), b.H.end(), cd(0, 0)); }
    kw.actions = 0;
    kw.max_res = 0;
    for (double& x : kw.ms) x = 0.0;
    const auto kin_t0 = std::chrono::steady_clock::now();
    ck(ctx->kry_w.ensure(mb));
    ck(ctx->kry_t1.ensure(mb));
    ck(ctx->small.ensure(64));
    ck(kw.tmp.ensure(20 * mb));
    const size_t batch_bytes = (size_t)o.maxnodes * mb;
    ck(kw.X.ensure(batch_bytes));
    ck(kw.A.ensure(batch_bytes));
    ck(kw.X2.ensure(batch_bytes));
    ck(kw.A2.ensure(batch_bytes));
    ck(kw.rhs2.ensure(2 * mb));
    ck(kw.LX.ensure(batch_bytes));
    ck(kw.T.ensure(batch_bytes));
    double* lev_c = kw.tmp.d() + 12 * mcount;    // current level covector / environment
    double* lev_e = kw.tmp.d() + 13 * mcount;
    double* c1 = kw.tmp.d() + 14 * mcount;       // one spec
    double* e1 = kw.tmp.d() + 15 * mcount;
    double* g = kw.tmp.d() + 16 * mcount;        // finished gradient of the previous level
    double* gnew = kw.tmp.d() + 17 * mcount;
    double* scr = kw.tmp.d() + 12 * mcount;      // overlaps nothing live during finish_gradient_dev? no: use extra
    (void)scr;
    DevBuf& fin = ctx->scratch;                  // 3 matrices for the finish
    ck(fin.ensure(12 * mb));
    auto level = [&](int nodes, double& val) {
        std::vector<double> theta, w;
        kinetic_grid(nodes, theta, w);
        val = 0;
        if (gradient) {
            ck(hipMemsetAsync(lev_c, 0, mb, ctx->stream));
            ck(hipMemsetAsync(lev_e, 0, mb, ctx->stream));
        }
        // the node solutions of every (op, direction) the level needs, their bases grown together
        std::vector<double> z(theta.size());
        for (size_t i = 0; i < theta.size(); ++i) z[i] = std::cos(theta[i]);
        double* Xs[2] = {kw.X.d(), kw.X2.d()};
        double* As[2] = {kw.A.d(), kw.A2.d()};
        std::vector<KinJob> jobs;
        for (int comm = 0; comm < 2; ++comm) {
            if (!(which & (1 << comm))) continue;
            const double* opd = comm ? st->Ch.d() : st->Rh.d();
            double* rhs = kw.rhs2.d() + (size_t)comm * mcount;
            dv.gemm(0, st->rho.d(), 0, opd, rhs);   // forward solves on rho op^+
            jobs.push_back(KinJob{&kw.basis[2 * comm], rhs, false, Xs[comm], 0, false});
            if (gradient) jobs.push_back(KinJob{&kw.basis[2 * comm + 1], opd, true, As[comm], 0, false});   // on op^+
        }
        kinetic_solutions(ctx, st, dv, jobs, z);
        for (int comm = 0; comm < 2; ++comm) {
            if (!(which & (1 << comm))) continue;
            double v = 0;
            kinetic_components(ctx, st, dv, comm, theta, w, gradient, v, c1, e1, Xs[comm], As[comm]);
            val += v;
            if (gradient) {
                dv.axpby(cd(1, 0), lev_c, cd(1, 0), c1, lev_c);
                dv.axpby(cd(1, 0), lev_e, cd(1, 0), e1, lev_e);
            }
        }
    };
    int nodes = o.minnodes;
    double prev = 0;
    level(nodes, prev);
    if (gradient) finish_gradient_dev(ctx, st, dv, lev_c, lev_e, metric_reg, g, fin.d(), S);
    while (2 * nodes <= o.maxnodes) {
        nodes *= 2;
        double cur = 0;
        level(nodes, cur);
        const bool energy_pass = std::fabs(cur - prev) <= o.atol + o.rtol * std::fabs(cur);
        bool done = false;
        if (gradient) {
            finish_gradient_dev(ctx, st, dv, lev_c, lev_e, metric_reg, gnew, fin.d(), S);
            dv.axpby(cd(1, 0), gnew, cd(-1, 0), g, c1);
            const double dg = dv.norm(c1), gn = dv.norm(gnew);
            if (!std::isfinite(cur) || !std::isfinite(dg) || !std::isfinite(gn)) throw Fail{RCMPS_ERR_NONFINITE};
            done = energy_pass && dg <= o.atol + o.rtol * gn;
            dv.copy(gnew, g);
        } else {
            if (!std::isfinite(cur)) throw Fail{RCMPS_ERR_NONFINITE};
            done = energy_pass;
        }
        prev = cur;
        if (done) {
            value = cur;
            if (gradient) {
                dv.copy(lev_c, cov);
                dv.copy(lev_e, env);
            }
            if (S) {
                S->kinetic_nodes = nodes;
                for (int i = 0; i < 4; ++i) S->kinetic_basis[i] = kw.basis[i].k;
                S->kinetic_actions = kw.actions;
                S->kinetic_residual = kw.max_res;
                ck(hipStreamSynchronize(ctx->stream));
                const double total =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - kin_t0).count();
                for (int i = 0; i < 3; ++i) S->kinetic_ms[i] = kw.ms[i];
                S->kinetic_ms[3] = total - kw.ms[0] - kw.ms[1] - kw.ms[2];
            }
            return;
        }
    }
    throw Fail{RCMPS_ERR_SOLVER};
}

bool valid_integrator(const rcmps_integrator* ip) {
    return ip && ip->n_quad >= 8 && ip->max_point > 0 && ip->atol > 0 && ip->rtol > 0 && ip->maxiters >= 1;
}

// A batch of evaluations: grouped by padded size, one persistent launch per group, then the
// pullback of each evaluation.  Every request gets its stats; the first failure is returned.
// The per-state pullback, boundary and metric solves of one class run on up to kFinishWorkers worker
// contexts, each driven by its own host thread and stream, while the main stream integrates the
// next class.  Each state is finished by one worker with the same kernels and tile choices as on the
// main context (every tile variant gives the same bits), so the results equal the sequential ones
// bit for bit.  Not for a single state or when a state appears twice (its cached metric inverse
// would be shared between threads).
constexpr int kFinishWorkers = 4;

bool can_finish_concurrently(const std::vector<int>& idx, const std::vector<EvalReq>& reqs) {
    if (idx.size() < 2) return false;
    std::vector<const rcmps_state*> seen;
    for (int i : idx) seen.push_back(reqs[i].st);
    std::sort(seen.begin(), seen.end());
    return std::adjacent_find(seen.begin(), seen.end()) == seen.end();
}

// Finish the evaluations idx (trajectories in ctx->tb[tb0 + j]) on the worker contexts.  Only this
// function touches ctx->workers; it may run on a host thread of its own while the main thread
// drives ctx->stream.
void finish_concurrently(rcmps_context* ctx, const rcmps_integrator* ip, double metric_reg, int flags,
                         const std::vector<int>& idx, size_t tb0, std::vector<EvalReq>& reqs) {
    const int nw = std::min<int>(kFinishWorkers, (int)idx.size());
    while ((int)ctx->workers.size() < nw) {
        rcmps_context* w = nullptr;
        const int rc = rcmps_context_create(ctx->device, &w);
        if (rc != RCMPS_OK) throw Fail{rc};
        w->gemm_force = ctx->gemm_force;
        ctx->workers.push_back(w);
    }
    for (int w = 0; w < nw; ++w) {
        ctx->workers[w]->gemm_force = ctx->gemm_force;
        ctx->workers[w]->layout = ctx->layout;
        ensure_quadrature(ctx->workers[w], ip);
    }
    std::atomic<int> next{0};
    auto work = [&](rcmps_context* wc) {
        (void)hipSetDevice(wc->device);
        for (;;) {
            const int j = next.fetch_add(1);
            if (j >= (int)idx.size()) break;
            EvalReq& r = reqs[idx[j]];
            if (r.S->status != RCMPS_OK || !r.adjoint()) continue;
            int code = RCMPS_OK;
            try {
                code = guard([&] {
                    finish_state(wc, r, *ctx->tb[tb0 + j], ip, metric_reg, flags);
                    ck(hipStreamSynchronize(wc->stream));
                });
            } catch (...) {   // nothing may leave a worker thread
                code = RCMPS_ERR_DEVICE;
            }
            if (code != RCMPS_OK) r.S->status = code;
        }
    };
    std::vector<std::thread> threads;
    for (int w = 1; w < nw; ++w) threads.emplace_back(work, ctx->workers[w]);
    work(ctx->workers[0]);
    for (auto& t : threads) t.join();
}

// The classes run smallest first: a class's finishing then runs under a longer launch, and only the
// last class's is left.  That one is cut in two launches, the second of about an eighth of its
// states, so that the first part finishes under the second launch.  Results do not depend on
// which launch integrates a state.  A class finished on the workers keeps its trajectory buffers
// until it is done, so consecutive launches alternate between two sets of buffers.
int fg_phi4_batch_impl(rcmps_context* ctx, std::vector<EvalReq>& reqs, const rcmps_integrator* ip,
                       double metric_reg, int flags) {
    if (!ctx || !valid_integrator(ip) || !(metric_reg >= 0)) return RCMPS_ERR_ARGUMENT;
    if (std::sqrt(ip->max_point) >= RC_SOURCE_TMAX) return RCMPS_ERR_UNSUPPORTED;
    const bool krylov = (flags & RCMPS_FLAG_KINETIC_KRYLOV) != 0;
    for (auto& r : reqs) {
        if (!r.st || !r.energy || !r.S || !std::isfinite(r.g)) return RCMPS_ERR_ARGUMENT;
        if (krylov && (r.grad_R || r.grad_K)) return RCMPS_ERR_ARGUMENT;
        if (r.st->Dp / 16 > 4) return RCMPS_ERR_UNSUPPORTED;
        std::memset(r.S, 0, sizeof(*r.S));
    }
    auto t_start = std::chrono::steady_clock::now();
    std::thread finisher;
    std::exception_ptr finish_error;
    auto join_finisher = [&] {
        if (finisher.joinable()) finisher.join();
        if (finish_error) {
            std::exception_ptr e = finish_error;
            finish_error = nullptr;
            std::rethrow_exception(e);
        }
    };
    int rc = guard([&] {
        try {
        ck(hipSetDevice(ctx->device));
        ensure_quadrature(ctx, ip);
        std::vector<std::vector<int>> classes;
        for (int Dp = 16; Dp <= 64; Dp += 16) {
            std::vector<int> idx;
            for (size_t i = 0; i < reqs.size(); ++i)
                if (reqs[i].st->Dp == Dp) idx.push_back((int)i);
            if (!idx.empty()) classes.push_back(std::move(idx));
        }
        if (!krylov && !classes.empty() && classes.back().size() >= 8) {
            std::vector<int>& last = classes.back();
            const size_t tail = (last.size() + 7) / 8;
            std::vector<int> second(last.end() - tail, last.end());
            last.resize(last.size() - tail);
            classes.push_back(std::move(second));
        }
        size_t largest = 0;
        for (const auto& idx : classes) largest = std::max(largest, idx.size());
        const bool overlap = !krylov && classes.size() > 1;
        // sized before any class starts: a worker thread reads ctx->tb while the next class runs
        while (ctx->tb.size() < (overlap ? 2 : 1) * largest) ctx->tb.emplace_back(new TrajBufs());
        for (size_t c = 0; c < classes.size(); ++c) {
            const std::vector<int>& idx = classes[c];
            const int Dp = reqs[idx[0]].st->Dp;
            const size_t tb0 = overlap ? (c & 1) * largest : 0;
            run_class(ctx, Dp, (flags & RCMPS_FLAG_SPLIT) != 0, krylov, (flags & RCMPS_FLAG_PHASE_TIMERS) != 0,
                      (flags & RCMPS_FLAG_DENSE) == 0, ip, idx, tb0, reqs);
            join_finisher();   // the previous class's finishing, overlapped with this class's launch
            if (!krylov && can_finish_concurrently(idx, reqs)) {
                finisher = std::thread([&, c, tb0] {
                    try {
                        finish_concurrently(ctx, ip, metric_reg, flags, classes[c], tb0, reqs);
                    } catch (...) {
                        finish_error = std::current_exception();
                    }
                });
                continue;
            }
            for (size_t j = 0; j < idx.size(); ++j) {
                EvalReq& r = reqs[idx[j]];
                if (r.S->status != RCMPS_OK) continue;
                const int code = guard([&] {
                    const double *kc = nullptr, *ke = nullptr;
                    if (krylov) {
                        // the kinetic job first, then the :phi^4: term (fixed order, observables.jl:236-246)
                        const size_t mcount = (size_t)Dp * Dp * 2;
                        ck(ctx->kin.tmp.ensure(20 * mcount * 8));
                        double* kcov = ctx->kin.tmp.d() + 18 * mcount;
                        double* kenv = ctx->kin.tmp.d() + 19 * mcount;
                        double kv = 0;
                        kinetic_resolvent_dev(ctx, r.st, 3, metric_reg, r.adjoint(), kv, kcov, kenv, r.S);
                        *r.energy = kv + *r.energy;
                        r.S->term_values[0] = kv;
                        r.S->term_values[1] = 0.0;
                        kc = kcov;
                        ke = kenv;
                    }
                    if (r.adjoint()) finish_state(ctx, r, *ctx->tb[tb0 + j], ip, metric_reg, flags, kc, ke);
                });
                if (code != RCMPS_OK) r.S->status = code;
            }
        }
        join_finisher();
        } catch (...) {
            if (finisher.joinable()) finisher.join();   // never leave a running thread behind
            throw;
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

int rcmps_set_layout(rcmps_context* ctx, int layout) {
    if (!ctx || (layout != RCMPS_LAYOUT_ROW_MAJOR && layout != RCMPS_LAYOUT_COLUMN_MAJOR)) return RCMPS_ERR_ARGUMENT;
    ctx->layout = layout;
    return RCMPS_OK;
}

int rcmps_set_nojump_threshold(rcmps_context* ctx, double kappa) {
    if (!ctx || !(kappa >= 0)) return RCMPS_ERR_ARGUMENT;
    ctx->nojump_kappa = kappa;
    return RCMPS_OK;
}

int rcmps_state_nojump_form(const rcmps_state* st, int* schur, double* kappa) {
    if (!st || !schur || !kappa) return RCMPS_ERR_ARGUMENT;
    *schur = st->nojump_schur;
    *kappa = st->nojump_kappa;
    return RCMPS_OK;
}

int rcmps_gemm_variant(rcmps_context* ctx, int variant) {
    if (!ctx || variant < -1 || variant >= kGemmVariants) return RCMPS_ERR_ARGUMENT;
    ctx->gemm_force = variant;
    return RCMPS_OK;
}

int rcmps_gemm_tuning(rcmps_context* ctx, int cap, int* padded_size, int* batch_class, int* variant, double* us,
                      int* count) {
    if (!ctx || !count || cap < 0 || (cap > 0 && (!padded_size || !batch_class || !variant || !us)))
        return RCMPS_ERR_ARGUMENT;
    int n = 0;
    for (const auto& kv : ctx->gemm_tuned) {
        if (n < cap) {
            padded_size[n] = kv.first.first;
            batch_class[n] = kv.first.second;
            variant[n] = kv.second.variant;
            for (int v = 0; v < kGemmVariants; ++v) us[kGemmVariants * n + v] = kv.second.us[v];
        }
        ++n;
    }
    *count = n;
    return RCMPS_OK;
}

int rcmps_set_kinetic_options(rcmps_context* ctx, const rcmps_kinetic_options* o) {
    if (!ctx || !o) return RCMPS_ERR_ARGUMENT;
    // src/kinetic.jl:78-82 and src/kinetic_krylov.jl:35-40
    if (!(o->minnodes >= 2 && o->maxnodes >= 2 * o->minnodes && std::isfinite(o->atol) && std::isfinite(o->rtol) &&
          o->atol > 0 && o->rtol >= 0 && std::isfinite(o->tolerance) && o->tolerance > 0 &&
          std::isfinite(o->fallback_tolerance) && o->fallback_tolerance >= o->tolerance && o->initialdim >= 1 &&
          o->maxdim >= o->initialdim))
        return RCMPS_ERR_ARGUMENT;
    ctx->kopt = *o;
    return RCMPS_OK;
}

int rcmps_kinetic(rcmps_context* ctx, const rcmps_state* st, int which, double metric_reg, double* value,
                  double* raw_cov, double* raw_env, double* grad, rcmps_stats* stats) {
    if (!ctx || !st || !value || which < 1 || which > 3 || !(metric_reg >= 0)) return RCMPS_ERR_ARGUMENT;
    rcmps_stats local{};
    rcmps_stats* S = stats ? stats : &local;
    std::memset(S, 0, sizeof(*S));
    auto t0 = std::chrono::steady_clock::now();
    const int rc = guard([&] {
        ck(hipSetDevice(ctx->device));
        const int D = st->D, Dp = st->Dp;
        Dev dv{ctx, Dp, D};
        const size_t mcount = (size_t)Dp * Dp * 2;
        ck(ctx->kin.tmp.ensure(20 * mcount * 8));
        double* kcov = ctx->kin.tmp.d() + 18 * mcount;
        double* kenv = ctx->kin.tmp.d() + 19 * mcount;
        const bool gradient = raw_cov || raw_env || grad;
        double v = 0;
        kinetic_resolvent_dev(ctx, st, which, metric_reg, gradient, v, kcov, kenv, S);
        *value = v;
        if (raw_cov) pad_out(ctx, D, Dp, kcov, raw_cov);
        if (raw_env) pad_out(ctx, D, Dp, kenv, raw_env);
        if (grad) {
            ck(ctx->scratch.ensure(12 * mcount * 8));
            double* G = ctx->kin.tmp.d() + 16 * mcount;
            finish_gradient_dev(ctx, st, dv, kcov, kenv, metric_reg, G, ctx->scratch.d(), S);
            pad_out(ctx, D, Dp, G, grad);
        }
        ck(hipStreamSynchronize(ctx->stream));
    });
    S->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    S->status = rc;
    return rc;
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
        const double* src = ctx->tb[ctx->last_tb]->samples[system].d();
        const int rows = ctx->last_npts * nl;
        for (int r = 0; r < rows; ++r)
            pad_out(ctx, D, Dp, src + (size_t)r * Dp * Dp * 2, dst + (size_t)r * D * D * 2);
        ck(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"

#include "excitation.inc"

// ==========================================================================================
// rcmps_boundary_environment_noalloc: the boundary Poisson solve of rcmps_boundary_environment on a
// fixed device array instead of allocated workspaces, for callers that cannot allocate.  Every
// call on one device shares that device's array: calls that overlap on one device race.
// ==========================================================================================
#define RC_NOALLOC_SCRATCH_BYTES ((size_t)4 << 20)
#define RC_NOALLOC_MAX_DEVICES 64

// Layout per call (mb = 16 Dp^2 bytes, m = min(D^2, 40) GMRES vectors): padded input and output
// (2 mb), t1, t2, z, b, x (5 mb), tmp (2 mb), the basis V ((m + 1) mb), w (mb), the Hessenberg column
// h, then 64 doubles of device scalars.  Everything a kernel reads is written earlier in the same
// call (pad_in and GMRES clear what has to start at zero), so stale contents never matter.
static __device__ __align__(256) double scratch[RC_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_ptr[RC_NOALLOC_MAX_DEVICES] = {};

hipError_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= RC_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    // the current device is `device` (rcmps_context_create set it): its copy of the array
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return hipSuccess;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// bytes of scratch one boundary solve at padded size Dp with m GMRES vectors takes
size_t noalloc_boundary_bytes(int Dp, int m) {
    const size_t mb = (size_t)Dp * Dp * 16;
    return (size_t)(2 + 5 + 2 + (m + 1) + 1) * mb + align256(gmres_col(m) * 8 + 64) + 64 * 8;
}

struct NoAllocScope {
    rcmps_context* ctx;
    explicit NoAllocScope(rcmps_context* c) : ctx(c) { ctx->no_alloc = true; }
    ~NoAllocScope() { ctx->no_alloc = false; }
};

}  // namespace

extern "C" int rcmps_boundary_environment_noalloc(rcmps_context* ctx, const rcmps_state* st, const double* a,
                                                  double tol, double* h, rcmps_stats* stats) {
    if (!ctx || !st || !a || !h || !(tol > 0)) return RCMPS_ERR_ARGUMENT;
    const int D = st->D, Dp = st->Dp;
    const int m = std::min(D * D, 40);
    if (noalloc_boundary_bytes(Dp, m) > RC_NOALLOC_SCRATCH_BYTES) return RCMPS_ERR_SCRATCH_LIMIT;
    double* base = (ctx->device >= 0 && ctx->device < RC_NOALLOC_MAX_DEVICES) ? g_scratch_ptr[ctx->device] : nullptr;
    if (!base) return RCMPS_ERR_DEVICE;
    // host read-back slots of the GMRES Hessenberg columns (pageable: nothing is allocated)
    double slots[2 * 166];
    static_assert(2 * 166 >= 2 * 4 * (40 + 1) + 4, "two GMRES columns at m = 40");
    NoAllocScope scope(ctx);
    return guard([&] {
        const size_t md = (size_t)Dp * Dp * 2;   // doubles per padded matrix
        double* p = base;
        double* ap = p; p += md;
        double* hp = p; p += md;
        BoundaryBufs bb{};
        bb.t1 = p; p += md;
        bb.t2 = p; p += md;
        bb.z = p; p += md;
        bb.b = p; p += md;
        bb.x = p; p += md;
        bb.tmp = p; p += 2 * md;
        bb.g.V = p; p += (size_t)(m + 1) * md;
        bb.g.w = p; p += md;
        bb.g.h = p; p += align256(gmres_col(m) * 8 + 64) / 8;
        bb.small = p;
        bb.g.host = slots;
        pad_in(ctx, D, Dp, a, ap);
        boundary_solve(ctx, st, ap, tol, tol, hp, stats, &bb);
        pad_out(ctx, D, Dp, hp, h);
        ck(hipStreamSynchronize(ctx->stream));
    });
}
