/*
 * FP64 CPU restatement of RcmpsCore's phi^4 energy-and-gradient path (fg_phi4).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked against: only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.  Nothing shipped
 * links or calls it.
 *
 * Conventions: complex128 matrices, row-major D x D (element (i,j) at i*D+j).
 * Citations are to the reference tree (AntoineTilloy/RcmpsCore.jl).
 *
 * Restated pieces:
 *   - Gauss nodes/weights on the mirrored window            src/observables.jl:65-75
 *   - source 2|t| J(t^2), J from K_{1/4}                    src/observables.jl:1-5
 *   - observable specs (channels, edges)                    src/observables.jl:7-29,255-256
 *   - literal ODE right-hand side (full channels)           src/observables.jl:37-63
 *   - Lindblad maps                                         src/transfer.jl:2-19
 *   - adaptive DOP853 with the reference GPU controller     gpu/src/rcmps_gpu/optimizer/runtime/
 *                                                           ode_backend.py:500-586,
 *                                                           device_control_backend.py:30-47
 *   - pullback (covector, environment)                      src/observables.jl:161-199
 *   - fixed-order linear combination                        src/observables.jl:226-254
 *   - finish_gradient / boundary Poisson / metric solve     src/observables.jl:201-212,
 *                                                           src/transfer.jl:122-153
 *   - NoJumpInverse (Schur + triangular Sylvester)          src/transfer.jl:50-68
 *   - CPTP Arnoldi stationary state, dense reference        src/transfer.jl:74-115
 *   - checked_density                                       src/transfer.jl:33-47
 *
 * The ODE of one (term, direction) pair is integrated in reduced coordinates: constant channels
 * (rho0 forward, I adjoint) are dropped, the terminal forward channel is replaced by its trace
 * (tr L(X) = 0 for a left-canonical state, so only the insertion terms feed it), and for the
 * kinetic specs channel 3 equals channel 2^dagger (forward and adjoint), so only channel 2 is
 * carried.  The derivative is still computed by the literal full-channel right-hand side of
 * src/observables.jl:37-63 on the unpacked channels; packing only selects entries.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "dop853_tableau.h"

typedef double _Complex cx;

#define OR_OK 0
#define OR_MAX_STEPS 1
#define OR_STEP_UNDERFLOW 2
#define OR_NONFINITE 3
#define OR_MISSING_SAMPLES 5
#define OR_SOLVER_FAILURE 10
#define OR_BAD_ARGUMENT 11

/* ------------------------------------------------------------------------------------------ */
/* dense helpers                                                                              */
/* ------------------------------------------------------------------------------------------ */

static cx* cx_alloc(size_t n) {
    cx* p = (cx*)calloc(n ? n : 1, sizeof(cx));
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* C = alpha * op(A) * op(B) + beta * C ; op in {0: none, 1: conjugate transpose}. No aliasing of C. */
static void mm(int n, cx alpha, const cx* A, int ta, const cx* B, int tb, cx beta, cx* C) {
    cx* b = cx_alloc((size_t)n * n);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            b[(size_t)k * n + j] = tb ? conj(B[(size_t)j * n + k]) : B[(size_t)k * n + j];
    if (beta == 0) {
        memset(C, 0, sizeof(cx) * (size_t)n * n);
    } else if (beta != 1) {
        for (size_t i = 0; i < (size_t)n * n; ++i) C[i] *= beta;
    }
    for (int i = 0; i < n; ++i) {
        cx* c = C + (size_t)i * n;
        for (int k = 0; k < n; ++k) {
            cx a = ta ? conj(A[(size_t)k * n + i]) : A[(size_t)i * n + k];
            a *= alpha;
            if (a == 0) continue;
            const cx* bk = b + (size_t)k * n;
            for (int j = 0; j < n; ++j) c[j] += a * bk[j];
        }
    }
    free(b);
}

static void adjoint_of(int n, const cx* A, cx* out) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[(size_t)i * n + j] = conj(A[(size_t)j * n + i]);
}

static cx trace_of(int n, const cx* A) {
    cx s = 0;
    for (int i = 0; i < n; ++i) s += A[(size_t)i * n + i];
    return s;
}

/* tr(A*B) */
static cx trace_prod(int n, const cx* A, const cx* B) {
    cx s = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * B[(size_t)k * n + i];
    return s;
}

static double fro_norm(size_t len, const cx* A) {
    double s = 0;
    for (size_t i = 0; i < len; ++i) s += creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]);
    return sqrt(s);
}

/* Solve A x = b for m right-hand sides (A is n x n, overwritten by its LU factors).
   B is n x m row-major, overwritten by the solution.  Partial pivoting. */
static int lu_solve(int n, cx* A, int m, cx* B) {
    int* piv = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = cabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = cabs(A[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (best == 0) { free(piv); return OR_SOLVER_FAILURE; }
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                cx t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)p * n + j]; A[(size_t)p * n + j] = t;
            }
            for (int j = 0; j < m; ++j) {
                cx t = B[(size_t)k * m + j]; B[(size_t)k * m + j] = B[(size_t)p * m + j]; B[(size_t)p * m + j] = t;
            }
        }
        cx inv = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            cx f = A[(size_t)i * n + k] * inv;
            if (f == 0) continue;
            A[(size_t)i * n + k] = f;
            cx* ai = A + (size_t)i * n;
            const cx* ak = A + (size_t)k * n;
            for (int j = k + 1; j < n; ++j) ai[j] -= f * ak[j];
            cx* bi = B + (size_t)i * m;
            const cx* bk = B + (size_t)k * m;
            for (int j = 0; j < m; ++j) bi[j] -= f * bk[j];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        cx inv = 1.0 / A[(size_t)k * n + k];
        cx* bk = B + (size_t)k * m;
        for (int i = k + 1; i < n; ++i) {
            cx f = A[(size_t)k * n + i];
            if (f == 0) continue;
            const cx* bi = B + (size_t)i * m;
            for (int j = 0; j < m; ++j) bk[j] -= f * bi[j];
        }
        for (int j = 0; j < m; ++j) bk[j] *= inv;
    }
    free(piv);
    return OR_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* complex Schur factorisation A = Z T Z^dagger (Householder Hessenberg + shifted QR)          */
/* ------------------------------------------------------------------------------------------ */

static void givens(cx a, cx b, double* c, cx* s) {
    /* [c s; -conj(s) c] [a; b] = [r; 0] */
    double aa = cabs(a), bb = cabs(b);
    if (bb == 0) { *c = 1; *s = 0; return; }
    if (aa == 0) { *c = 0; *s = conj(b) / bb; return; }
    double nrm = hypot(aa, bb);
    *c = aa / nrm;
    *s = (a / aa) * conj(b) / nrm;
}

int or_schur(int n, cx* A, cx* Z) {
    /* A (n x n) -> T upper triangular in place; Z unitary with A_in = Z T Z^dagger. */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[(size_t)i * n + j] = (i == j);
    cx* v = cx_alloc((size_t)n);
    for (int k = 0; k + 2 < n + 0 && k < n - 2; ++k) {
        double alpha = 0;
        for (int i = k + 1; i < n; ++i) alpha += creal(A[(size_t)i * n + k] * conj(A[(size_t)i * n + k]));
        alpha = sqrt(alpha);
        if (alpha == 0) continue;
        cx x0 = A[(size_t)(k + 1) * n + k];
        cx phase = cabs(x0) > 0 ? x0 / cabs(x0) : 1.0;
        for (int i = 0; i < n; ++i) v[i] = 0;
        for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
        v[k + 1] += phase * alpha;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += creal(v[i] * conj(v[i]));
        if (vn == 0) continue;
        /* H = I - 2 v v^dagger / vn ; A <- H A H ; Z <- Z H */
        for (int j = 0; j < n; ++j) {
            cx s = 0;
            for (int i = k + 1; i < n; ++i) s += conj(v[i]) * A[(size_t)i * n + j];
            s *= 2.0 / vn;
            for (int i = k + 1; i < n; ++i) A[(size_t)i * n + j] -= v[i] * s;
        }
        for (int i = 0; i < n; ++i) {
            cx s = 0;
            for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
            s *= 2.0 / vn;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= s * conj(v[j]);
        }
        for (int i = 0; i < n; ++i) {
            cx s = 0;
            for (int j = k + 1; j < n; ++j) s += Z[(size_t)i * n + j] * v[j];
            s *= 2.0 / vn;
            for (int j = k + 1; j < n; ++j) Z[(size_t)i * n + j] -= s * conj(v[j]);
        }
        for (int i = k + 2; i < n; ++i) A[(size_t)i * n + k] = 0;
    }
    free(v);
    /* shifted QR on the active window [lo, hi] */
    int hi = n - 1, iter = 0, total = 0;
    const double eps = 2.220446049250313e-16;
    while (hi > 0) {
        int lo = hi;
        while (lo > 0) {
            double s = cabs(A[(size_t)(lo - 1) * n + lo - 1]) + cabs(A[(size_t)lo * n + lo]);
            if (s == 0) s = 1;
            if (cabs(A[(size_t)lo * n + lo - 1]) <= eps * s) { A[(size_t)lo * n + lo - 1] = 0; break; }
            --lo;
        }
        if (lo == hi) { --hi; iter = 0; continue; }
        if (++total > 100 * n + 1000) return OR_SOLVER_FAILURE;
        ++iter;
        /* Wilkinson shift from the trailing 2x2 block */
        cx a = A[(size_t)(hi - 1) * n + hi - 1], b = A[(size_t)(hi - 1) * n + hi];
        cx c = A[(size_t)hi * n + hi - 1], d = A[(size_t)hi * n + hi];
        cx tr = a + d, det = a * d - b * c;
        cx disc = csqrt(tr * tr / 4.0 - det);
        cx mu1 = tr / 2.0 + disc, mu2 = tr / 2.0 - disc;
        cx mu = cabs(mu1 - d) < cabs(mu2 - d) ? mu1 : mu2;
        if (iter % 11 == 10) mu = d + cabs(creal(A[(size_t)hi * n + hi - 1])) + 0.75 * cabs(A[(size_t)hi * n + hi - 1]);
        /* implicit single-shift QR sweep via Givens rotations */
        cx x = A[(size_t)lo * n + lo] - mu, y = A[(size_t)(lo + 1) * n + lo];
        for (int k = lo; k < hi; ++k) {
            double cg; cx sg;
            if (k > lo) { x = A[(size_t)k * n + k - 1]; y = A[(size_t)(k + 1) * n + k - 1]; }
            givens(x, y, &cg, &sg);
            /* rows k, k+1 */
            for (int j = (k > lo ? k - 1 : lo); j < n; ++j) {
                cx t1 = A[(size_t)k * n + j], t2 = A[(size_t)(k + 1) * n + j];
                A[(size_t)k * n + j] = cg * t1 + sg * t2;
                A[(size_t)(k + 1) * n + j] = -conj(sg) * t1 + cg * t2;
            }
            /* columns k, k+1 */
            int imax = (k + 2 < hi ? k + 2 : hi);
            for (int i = 0; i <= imax; ++i) {
                cx t1 = A[(size_t)i * n + k], t2 = A[(size_t)i * n + k + 1];
                A[(size_t)i * n + k] = cg * t1 + conj(sg) * t2;
                A[(size_t)i * n + k + 1] = -sg * t1 + cg * t2;
            }
            for (int i = 0; i < n; ++i) {
                cx t1 = Z[(size_t)i * n + k], t2 = Z[(size_t)i * n + k + 1];
                Z[(size_t)i * n + k] = cg * t1 + conj(sg) * t2;
                Z[(size_t)i * n + k + 1] = -sg * t1 + cg * t2;
            }
            if (k > lo) A[(size_t)(k + 1) * n + k - 1] = 0;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) A[(size_t)i * n + j] = 0;
    return OR_OK;
}

/* Solve T X + X T^dagger = C (adjoint=0) or T^dagger X + X T = C (adjoint=1), T upper
   triangular.  This is LAPACK trsyl('N','C') / trsyl('C','N') with scale 1. */
static int trsyl_upper(int n, const cx* T, cx* C, int adjoint) {
    if (!adjoint) {
        for (int i = n - 1; i >= 0; --i)
            for (int j = n - 1; j >= 0; --j) {
                cx s = C[(size_t)i * n + j];
                for (int k = i + 1; k < n; ++k) s -= T[(size_t)i * n + k] * C[(size_t)k * n + j];
                for (int k = j + 1; k < n; ++k) s -= C[(size_t)i * n + k] * conj(T[(size_t)j * n + k]);
                cx den = T[(size_t)i * n + i] + conj(T[(size_t)j * n + j]);
                if (den == 0) return OR_SOLVER_FAILURE;
                C[(size_t)i * n + j] = s / den;
            }
    } else {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                cx s = C[(size_t)i * n + j];
                for (int k = 0; k < i; ++k) s -= conj(T[(size_t)k * n + i]) * C[(size_t)k * n + j];
                for (int k = 0; k < j; ++k) s -= C[(size_t)i * n + k] * T[(size_t)k * n + j];
                cx den = conj(T[(size_t)i * n + i]) + T[(size_t)j * n + j];
                if (den == 0) return OR_SOLVER_FAILURE;
                C[(size_t)i * n + j] = s / den;
            }
    }
    return OR_OK;
}

/* spectral norm by power iteration on A^dagger A (sufficient for the preconditioner shift) */
static double opnorm2(int n, const cx* A) {
    cx* x = cx_alloc((size_t)n);
    cx* y = cx_alloc((size_t)n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * i;
    double lam = 0;
    for (int it = 0; it < 300; ++it) {
        for (int i = 0; i < n; ++i) {
            cx s = 0;
            for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * x[k];
            y[i] = s;
        }
        for (int k = 0; k < n; ++k) {
            cx s = 0;
            for (int i = 0; i < n; ++i) s += conj(A[(size_t)i * n + k]) * y[i];
            x[k] = s;
        }
        double nx = fro_norm((size_t)n, x);
        if (nx == 0) break;
        double prev = lam;
        lam = sqrt(nx);
        for (int k = 0; k < n; ++k) x[k] /= nx;
        if (it > 10 && fabs(lam - prev) <= 1e-14 * lam) break;
    }
    free(x); free(y);
    return lam;
}

typedef struct {
    int n;
    cx* T;   /* Schur form minus shift/2 */
    cx* U;   /* Schur vectors */
    double shift;
} nojump_t;

/* src/transfer.jl:50-60 */
static int nojump_init(nojump_t* p, int n, const cx* q, const cx* r) {
    p->n = n;
    p->T = cx_alloc((size_t)n * n);
    p->U = cx_alloc((size_t)n * n);
    memcpy(p->T, q, sizeof(cx) * (size_t)n * n);
    int st = or_schur(n, p->T, p->U);
    if (st) return st;
    double nq = opnorm2(n, q), nr = opnorm2(n, r);
    double scale = fmax(1.0, fmax(nq, nr * nr));
    double maxre = -INFINITY;
    for (int i = 0; i < n; ++i) maxre = fmax(maxre, creal(p->T[(size_t)i * n + i]));
    p->shift = maxre >= -100 * 2.220446049250313e-16 * scale ? scale : 0.0;
    for (int i = 0; i < n; ++i) p->T[(size_t)i * n + i] -= p->shift / 2;
    return OR_OK;
}

static void nojump_free(nojump_t* p) { free(p->T); free(p->U); }

/* src/transfer.jl:61-68: out = U (-X) U^dagger with X solving the Sylvester equation in U basis */
static int nojump_apply(const nojump_t* p, const cx* y, cx* out, int adjoint) {
    int n = p->n;
    cx* t = cx_alloc((size_t)n * n);
    cx* c = cx_alloc((size_t)n * n);
    mm(n, 1, p->U, 1, y, 0, 0, t);
    mm(n, 1, t, 0, p->U, 0, 0, c);
    int st = trsyl_upper(n, p->T, c, adjoint);
    if (!st) {
        for (size_t i = 0; i < (size_t)n * n; ++i) c[i] = -c[i];
        mm(n, 1, p->U, 0, c, 0, 0, t);
        mm(n, 1, t, 0, p->U, 1, 0, out);
    }
    free(t); free(c);
    return st;
}

/* ------------------------------------------------------------------------------------------ */
/* Lindblad maps (src/transfer.jl:2-19)                                                       */
/* ------------------------------------------------------------------------------------------ */

static void lindblad(int n, const cx* q, const cx* r, const cx* x, cx* out, int adjoint) {
    cx* tmp = cx_alloc((size_t)n * n);
    if (!adjoint) {
        mm(n, 1, q, 0, x, 0, 0, out);
        mm(n, 1, x, 0, q, 1, 1, out);
        mm(n, 1, r, 0, x, 0, 0, tmp);
        mm(n, 1, tmp, 0, r, 1, 1, out);
    } else {
        mm(n, 1, q, 1, x, 0, 0, out);
        mm(n, 1, x, 0, q, 0, 1, out);
        mm(n, 1, r, 1, x, 0, 0, tmp);
        mm(n, 1, tmp, 0, r, 0, 1, out);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------------------------ */
/* quadrature (src/observables.jl:65-75) and source (src/observables.jl:1-5)                  */
/* ------------------------------------------------------------------------------------------ */

/* Gauss-Legendre nodes on (-1,1) by Newton iteration on P_n; weights 2/((1-x^2) P_n'(x)^2). */
static void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            double pn = n == 0 ? 1 : p1, pm = p0;
            if (n == 1) { pn = z; pm = 1; }
            pp = n * (z * pn - pm) / (z * z - 1);
            double dz = pn / pp;
            z -= dz;
            if (fabs(dz) < 1e-17) break;
        }
        {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            pp = n * (z * p1 - p0) / (z * z - 1);
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1 - z * z) * pp * pp);
    }
}

/* points/weights of length 2n: mirrored Gauss nodes of (0, sqrt(max_point)) */
void or_quadrature(int n, double max_point, double* points, double* weights) {
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    gauss_legendre(n, x, w);
    double a = sqrt(max_point);
    for (int i = 0; i < n; ++i) {
        double p = a / 2 * (x[i] + 1), wt = a / 2 * w[i];
        points[n + i] = p;
        weights[n + i] = wt;
        points[n - 1 - i] = -p;
        weights[n - 1 - i] = wt;
    }
    free(x); free(w);
}

/* K_{1/4}(x) by the trapezoid rule on K_nu(x) = int_0^inf exp(-x cosh u) cosh(nu u) du,
   exponentially convergent for this analytic, double-exponentially decaying integrand. */
static double bessel_k_quarter(double x) {
    const double h = 1.0 / 64;
    double s = 0.5 * exp(-x);
    for (int k = 1; k < 200000; ++k) {
        double u = k * h;
        double term = exp(-x * cosh(u)) * cosh(0.25 * u);
        s += term;
        if (term < 1e-19 * s && x * cosh(u) > 50) break;
    }
    return s * h;
}

/* transformed_source(t) = 2|t| J(t^2) with J(x) = K_{1/4}(|x|) / (2^{9/4} sqrt(pi) Gamma(5/4) |x|^{1/4}) */
double or_source(double t) {
    double a = fabs(t);
    if (a == 0) return 1.0 / sqrt(M_PI);
    double x = a * a;
    double c = pow(2.0, 2.25) * sqrt(M_PI) * tgamma(1.25);
    return 2 * a * bessel_k_quarter(x) / (c * pow(x, 0.25));
}

/* ------------------------------------------------------------------------------------------ */
/* observable specs and literal RHS (src/observables.jl:7-63)                                 */
/* ------------------------------------------------------------------------------------------ */

typedef struct { int src, dst, left; double coefficient; } edge_t;
typedef struct { int channels, nedges, commutator; edge_t edges[16]; } spec_t;

static spec_t kinetic_spec(int commutator) {
    spec_t s;
    memset(&s, 0, sizeof s);
    s.channels = 4; s.commutator = commutator; s.nedges = 4;
    double c = sqrt(2.0);
    s.edges[0] = (edge_t){0, 1, 1, c};
    s.edges[1] = (edge_t){0, 2, 0, c};
    s.edges[2] = (edge_t){2, 3, 1, c};
    s.edges[3] = (edge_t){1, 3, 0, c};
    return s;
}

static spec_t monomial_spec(int n) {
    spec_t s;
    memset(&s, 0, sizeof s);
    s.channels = n + 1; s.commutator = 0; s.nedges = 0;
    for (int k = 0; k < n; ++k) {
        s.edges[s.nedges++] = (edge_t){k, k + 1, 1, (double)(k + 1)};
        s.edges[s.nedges++] = (edge_t){k, k + 1, 0, (double)(k + 1)};
    }
    return s;
}

typedef struct {
    int n;
    const cx *q, *r, *op;   /* op = R or [Q,R] */
    spec_t spec;
    int adj;
} rhs_ctx_t;

/* du[k] for all channels from u[k]; src/observables.jl:37-63 */
static void rhs_full(const rhs_ctx_t* c, double t, const cx* u, cx* du) {
    int n = c->n;
    size_t m = (size_t)n * n;
    double jac = 2 * fabs(t), j = or_source(t);
    for (int k = 0; k < c->spec.channels; ++k) {
        lindblad(n, c->q, c->r, u + k * m, du + k * m, c->adj);
        for (size_t i = 0; i < m; ++i) du[k * m + i] *= jac;
    }
    for (int e = 0; e < c->spec.nedges; ++e) {
        edge_t ed = c->spec.edges[e];
        cx a = ed.coefficient * j;
        if (c->adj) {
            if (ed.left) mm(n, a, u + ed.dst * m, 0, c->op, 0, 1, du + ed.src * m);
            else mm(n, a, c->op, 1, u + ed.dst * m, 0, 1, du + ed.src * m);
        } else {
            if (ed.left) mm(n, a, c->op, 0, u + ed.src * m, 0, 1, du + ed.dst * m);
            else mm(n, a, u + ed.src * m, 0, c->op, 1, 1, du + ed.dst * m);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* reduced systems                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* term 0: kinetic (op R), term 1: kinetic (op [Q,R]), term 2: :phi^4: (op R)
   system index = 2*term + adj */
typedef struct {
    rhs_ctx_t ctx;
    int term, adj;
    const cx* rho;
    int nlive;           /* live matrix channels */
    int live[8];         /* spec channel index of each live matrix */
    int has_scalar;      /* forward: trace of terminal channel */
    int mirror_from;     /* kinetic: channel 2 = channel 1 ^dagger; -1 if none */
    size_t len;          /* complex entries of the reduced state */
} system_t;

static void system_init(system_t* s, int n, const cx* q, const cx* r, const cx* comm, const cx* rho,
                        int term, int adj) {
    memset(s, 0, sizeof *s);
    s->ctx.n = n; s->ctx.q = q; s->ctx.r = r; s->ctx.adj = adj;
    s->ctx.op = term == 1 ? comm : r;
    s->ctx.spec = term == 2 ? monomial_spec(4) : kinetic_spec(term == 1);
    s->term = term; s->adj = adj; s->rho = rho;
    s->mirror_from = -1;
    if (term < 2) {
        if (!adj) { s->nlive = 1; s->live[0] = 1; s->has_scalar = 1; }
        else { s->nlive = 2; s->live[0] = 0; s->live[1] = 1; }
        s->mirror_from = 1;
    } else {
        if (!adj) { s->nlive = 3; s->live[0] = 1; s->live[1] = 2; s->live[2] = 3; s->has_scalar = 1; }
        else { s->nlive = 4; for (int k = 0; k < 4; ++k) s->live[k] = k; }
    }
    s->len = (size_t)s->nlive * n * n + (size_t)s->has_scalar;
}

/* unpack reduced y into the full channel array (src/observables.jl trajectory layout) */
static void system_unpack(const system_t* s, const cx* y, cx* full) {
    int n = s->ctx.n;
    size_t m = (size_t)n * n;
    int ch = s->ctx.spec.channels;
    memset(full, 0, sizeof(cx) * m * ch);
    if (!s->adj) memcpy(full, s->rho, sizeof(cx) * m);
    else for (int i = 0; i < n; ++i) full[(size_t)(ch - 1) * m + (size_t)i * n + i] = 1;
    for (int k = 0; k < s->nlive; ++k) memcpy(full + s->live[k] * m, y + k * m, sizeof(cx) * m);
    if (s->mirror_from >= 0) adjoint_of(n, full + s->mirror_from * m, full + (s->mirror_from + 1) * m);
    /* terminal forward channel left at zero: only its trace is carried */
}

typedef struct { system_t* sys; cx* full; cx* dfull; long calls; } sysrhs_t;

static void system_rhs(sysrhs_t* w, double t, const cx* y, cx* dy) {
    const system_t* s = w->sys;
    int n = s->ctx.n;
    size_t m = (size_t)n * n;
    int ch = s->ctx.spec.channels;
    system_unpack(s, y, w->full);
    rhs_full(&s->ctx, t, w->full, w->dfull);
    for (int k = 0; k < s->nlive; ++k) memcpy(dy + k * m, w->dfull + s->live[k] * m, sizeof(cx) * m);
    if (s->has_scalar) dy[(size_t)s->nlive * m] = trace_of(n, w->dfull + (size_t)(ch - 1) * m);
    w->calls++;
}

/* ------------------------------------------------------------------------------------------ */
/* DOP853 (ode_backend.py:500-586; controller device_control_backend.py:30-47)                 */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    long accepted, rejected, attempts, dense_steps, rhs_calls;
    int status;
} traj_stats_t;

static int integrate(sysrhs_t* w, const double* points, int npoints, double atol, double rtol,
                     double initial_step, long max_steps, cx* samples, traj_stats_t* st) {
    size_t len = w->sys->len;
    double a = fabs(points[npoints - 1]);
    cx* y = cx_alloc(len);
    cx* ynew = cx_alloc(len);
    cx* tmp = cx_alloc(len);
    cx* f = cx_alloc(len);
    cx* K = cx_alloc(len * OR_N_STAGES_EXT);
    cx* dense = cx_alloc(len * OR_INTERP_POWER);
    memset(st, 0, sizeof *st);
    /* initial state: forward channels start at zero (rho0 is constant), adjoint channels at zero
       except the constant identity; the scalar starts at zero */
    double t = -a, h = fmin(initial_step > 0 ? initial_step : 0.01, 2 * a);
    system_rhs(w, t, y, f);
    int next_point = 0, rejected_since_accept = 0, status = OR_OK;
    while (t < a) {
        double endpoint = t < 0 ? 0.0 : a;
        h = fmin(h, endpoint - t);
        if (h <= 10 * fabs(nextafter(t, INFINITY) - t)) { status = OR_STEP_UNDERFLOW; break; }
        if (st->attempts >= max_steps) { status = OR_MAX_STEPS; break; }
        st->attempts++;
        memcpy(K, f, sizeof(cx) * len);
        for (int s = 1; s < OR_N_STAGES; ++s) {
            for (size_t i = 0; i < len; ++i) {
                cx acc = 0;
                for (int j = 0; j < s; ++j) acc += OR_A[s][j] * K[j * len + i];
                tmp[i] = y[i] + h * acc;
            }
            system_rhs(w, t + OR_C[s] * h, tmp, K + s * len);
        }
        for (size_t i = 0; i < len; ++i) {
            cx acc = 0;
            for (int j = 0; j < OR_N_STAGES; ++j) acc += OR_B[j] * K[j * len + i];
            ynew[i] = y[i] + h * acc;
        }
        system_rhs(w, t + h, ynew, K + OR_N_STAGES * len);
        double e5 = 0, e3 = 0;
        for (size_t i = 0; i < len; ++i) {
            double sc = atol + rtol * fmax(cabs(y[i]), cabs(ynew[i]));
            cx a5 = 0, a3 = 0;
            for (int j = 0; j <= OR_N_STAGES; ++j) {
                a5 += OR_E5[j] * K[j * len + i];
                a3 += OR_E3[j] * K[j * len + i];
            }
            a5 /= sc; a3 /= sc;
            e5 += creal(a5) * creal(a5) + cimag(a5) * cimag(a5);
            e3 += creal(a3) * creal(a3) + cimag(a3) * cimag(a3);
        }
        double den = e5 + 0.01 * e3;
        double error = den > 0 ? h * e5 / sqrt(den * (double)len) : 0.0;
        if (!isfinite(error)) { status = OR_NONFINITE; break; }
        double factor = error > 0 ? fmin(10.0, 0.9 * pow(error, -1.0 / 8.0)) : 10.0;
        if (error > 1) {
            st->rejected++;
            rejected_since_accept = 1;
            h *= fmax(0.2, factor);
            continue;
        }
        st->accepted++;
        double t_new = h == endpoint - t ? endpoint : t + h;
        if (next_point < npoints && points[next_point] <= t_new) {
            st->dense_steps++;
            for (int s = OR_N_STAGES + 1; s < OR_N_STAGES_EXT; ++s) {
                for (size_t i = 0; i < len; ++i) {
                    cx acc = 0;
                    for (int j = 0; j < s; ++j) acc += OR_A[s][j] * K[j * len + i];
                    tmp[i] = y[i] + h * acc;
                }
                system_rhs(w, t + OR_C[s] * h, tmp, K + s * len);
            }
            for (size_t i = 0; i < len; ++i) {
                cx delta = ynew[i] - y[i];
                dense[i] = delta;
                dense[len + i] = h * f[i] - delta;
                dense[2 * len + i] = 2.0 * delta - h * (K[OR_N_STAGES * len + i] + f[i]);
                for (int r = 0; r < 4; ++r) {
                    cx acc = 0;
                    for (int j = 0; j < OR_N_STAGES_EXT; ++j) acc += OR_D[r][j] * K[j * len + i];
                    dense[(3 + r) * len + i] = h * acc;
                }
            }
            while (next_point < npoints && points[next_point] <= t_new) {
                double p = points[next_point];
                cx* out = samples + (size_t)next_point * len;
                if (p == t_new) {
                    memcpy(out, ynew, sizeof(cx) * len);
                } else {
                    double x = (p - t) / h;
                    for (size_t i = 0; i < len; ++i) {
                        cx res = 0;
                        for (int r = OR_INTERP_POWER - 1, k = 0; r >= 0; --r, ++k) {
                            res += dense[(size_t)r * len + i];
                            res *= (k % 2 == 0) ? x : 1 - x;
                        }
                        out[i] = res + y[i];
                    }
                }
                next_point++;
            }
        }
        memcpy(y, ynew, sizeof(cx) * len);
        memcpy(f, K + OR_N_STAGES * len, sizeof(cx) * len);
        t = t_new;
        h *= rejected_since_accept ? fmin(1.0, factor) : factor;
        rejected_since_accept = 0;
    }
    if (status == OR_OK && next_point != npoints) status = OR_MISSING_SAMPLES;
    for (size_t i = 0; i < len * (size_t)npoints && status == OR_OK; ++i)
        if (!isfinite(creal(samples[i])) || !isfinite(cimag(samples[i]))) status = OR_NONFINITE;
    st->status = status;
    st->rhs_calls = w->calls;
    free(y); free(ynew); free(tmp); free(f); free(K); free(dense);
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* pullback (src/observables.jl:161-199), on unpacked full channels                           */
/* ------------------------------------------------------------------------------------------ */

static void addtriple(int n, cx* out, const cx* a, int ta, const cx* b, int tb, const cx* c, int tc,
                      cx coefficient, cx* tmp) {
    mm(n, 1, a, ta, b, tb, 0, tmp);
    mm(n, coefficient, tmp, 0, c, tc, 1, out);
}

/* covector (horizontal directions dQ = -R^+ dR) and environment; when xr, xq are given, also the
 * separate coefficients of dR and dQ, dvalue = 2 Re tr(xr dR) + 2 Re tr(xq dQ): M and N split as
 * M = MR - MQ R^+, N = NR - R NQ with MR = sum w x R^+ a, MQ = sum w x a, NR = sum w a R x,
 * NQ = sum w a x, and xr = (MR + NR^+)/2, xq = (MQ + NQ^+)/2 plus the edge-operator terms. */
static void pullback(const system_t* fw, const system_t* bw, const cx* fsamples, const cx* bsamples,
                     const double* points, const double* weights, int nnodes,
                     cx* covector, cx* env, double* value, cx* xr, cx* xq) {
    int n = fw->ctx.n;
    size_t m = (size_t)n * n;
    const cx *q = fw->ctx.q, *r = fw->ctx.r;
    spec_t spec = fw->ctx.spec;
    int ch = spec.channels;
    cx *M = cx_alloc(m), *N = cx_alloc(m), *U = cx_alloc(m), *V = cx_alloc(m), *tmp = cx_alloc(m);
    cx *MR = cx_alloc(m), *MQ = cx_alloc(m), *NR = cx_alloc(m), *NQ = cx_alloc(m);
    cx* rho = cx_alloc(m * ch);
    cx* obs = cx_alloc(m * ch);
    for (int i = 0; i < nnodes; ++i) {
        system_unpack(fw, fsamples + (size_t)i * fw->len, rho);
        system_unpack(bw, bsamples + (size_t)(nnodes - 1 - i) * bw->len, obs);
        double w = weights[i] * 2 * fabs(points[i]);
        for (int k = 0; k < ch; ++k) {
            const cx *x = rho + k * m, *a = obs + k * m;
            addtriple(n, M, x, 0, r, 1, a, 0, w, tmp);
            addtriple(n, M, x, 0, a, 0, r, 1, -w, tmp);
            addtriple(n, N, a, 0, r, 0, x, 0, w, tmp);
            addtriple(n, N, r, 0, a, 0, x, 0, -w, tmp);
            if (xr) {
                addtriple(n, MR, x, 0, r, 1, a, 0, w, tmp);
                mm(n, w, x, 0, a, 0, 1, MQ);
                addtriple(n, NR, a, 0, r, 0, x, 0, w, tmp);
                mm(n, w, a, 0, x, 0, 1, NQ);
            }
        }
        double j = weights[i] * or_source(points[i]);
        for (int e = 0; e < spec.nedges; ++e) {
            edge_t ed = spec.edges[e];
            if (ed.left) mm(n, j * ed.coefficient, rho + ed.src * m, 0, obs + ed.dst * m, 0, 1, U);
            else mm(n, j * ed.coefficient, obs + ed.dst * m, 0, rho + ed.src * m, 0, 1, V);
        }
    }
    if (xr) {
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                size_t ik = (size_t)i * n + k, ki = (size_t)k * n + i;
                xr[ik] = (MR[ik] + conj(NR[ki])) / 2;
                xq[ik] = (MQ[ik] + conj(NQ[ki])) / 2;
            }
        /* edge operator O with coefficient U + V^+ (left and right insertions) */
        cx* W = cx_alloc(m);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) W[(size_t)i * n + k] = (U[(size_t)i * n + k] + conj(V[(size_t)k * n + i])) / 2;
        if (spec.commutator) {   /* O = QR - RQ: dO = dQ R + Q dR - dR Q - R dQ */
            mm(n, 1, W, 0, q, 0, 1, xr); mm(n, -1, q, 0, W, 0, 1, xr);
            mm(n, 1, r, 0, W, 0, 1, xq); mm(n, -1, W, 0, r, 0, 1, xq);
        } else {
            for (size_t i = 0; i < m; ++i) xr[i] += W[i];
        }
        free(W);
    }
    if (spec.commutator) {
        mm(n, 1, U, 0, q, 0, 1, M); mm(n, -1, q, 0, U, 0, 1, M);
        addtriple(n, M, r, 0, U, 0, r, 1, -1, tmp);
        addtriple(n, M, U, 0, r, 0, r, 1, 1, tmp);
        mm(n, 1, q, 1, V, 0, 1, N); mm(n, -1, V, 0, q, 1, 1, N);
        addtriple(n, N, r, 0, V, 0, r, 1, -1, tmp);
        addtriple(n, N, r, 0, r, 1, V, 0, 1, tmp);
    } else {
        for (size_t i = 0; i < m; ++i) { M[i] += U[i]; N[i] += V[i]; }
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            covector[(size_t)i * n + k] = (M[(size_t)i * n + k] + conj(N[(size_t)k * n + i])) / 2;
    /* environment: adjoint channel 1 at the endpoint, hermitised */
    system_unpack(bw, bsamples + (size_t)nnodes * bw->len, obs);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            env[(size_t)i * n + k] = (obs[(size_t)i * n + k] + conj(obs[(size_t)k * n + i])) / 2;
    /* value: tr of the terminal forward channel = carried scalar */
    *value = creal(fsamples[(size_t)nnodes * fw->len + fw->len - 1]);
    free(M); free(N); free(U); free(V); free(tmp); free(rho); free(obs);
    free(MR); free(MQ); free(NR); free(NQ);
}

/* ------------------------------------------------------------------------------------------ */
/* boundary Poisson equation (src/transfer.jl:122-153) and metric solve                       */
/* ------------------------------------------------------------------------------------------ */

/* op(h) = -L*(h) + tr(rho h) I */
static void boundary_op(int n, const cx* q, const cx* r, const cx* rho, const cx* h, cx* out) {
    lindblad(n, q, r, h, out, 1);
    cx t = trace_prod(n, rho, h);
    for (size_t i = 0; i < (size_t)n * n; ++i) out[i] = -out[i];
    for (int i = 0; i < n; ++i) out[(size_t)i * n + i] += t;
}

/* dense solve of a linear operator given as a function on D x D matrices */
typedef void (*matop_fn)(void* ctx, const cx* x, cx* y);

static int dense_operator_solve(int n, matop_fn f, void* ctx, const cx* b, cx* x) {
    size_t m = (size_t)n * n;
    cx* A = cx_alloc(m * m);
    cx* e = cx_alloc(m);
    cx* col = cx_alloc(m);
    for (size_t j = 0; j < m; ++j) {
        memset(e, 0, sizeof(cx) * m);
        e[j] = 1;
        f(ctx, e, col);
        for (size_t i = 0; i < m; ++i) A[i * m + j] = col[i];
    }
    memcpy(x, b, sizeof(cx) * m);
    int st = lu_solve((int)m, A, 1, x);
    free(A); free(e); free(col);
    return st;
}

typedef struct { int n; const cx *q, *r, *rho; const nojump_t* p; const cx* z; cx den; cx* tmp; } bnd_ctx_t;

static void bnd_op(void* c, const cx* x, cx* y) {
    bnd_ctx_t* b = (bnd_ctx_t*)c;
    boundary_op(b->n, b->q, b->r, b->rho, x, y);
}

/* Sherman-Morrison right preconditioner: x = P*(y); x - tr(rho x)/den z */
static void bnd_pre(bnd_ctx_t* b, const cx* y, cx* x) {
    nojump_apply(b->p, y, x, 1);
    cx t = trace_prod(b->n, b->rho, x) / b->den;
    for (size_t i = 0; i < (size_t)b->n * b->n; ++i) x[i] -= t * b->z[i];
}

static void bnd_op_pre(void* c, const cx* x, cx* y) {
    bnd_ctx_t* b = (bnd_ctx_t*)c;
    bnd_pre(b, x, b->tmp);
    boundary_op(b->n, b->q, b->r, b->rho, b->tmp, y);
}

/* restarted GMRES(krylovdim) with classical Gram-Schmidt twice; Givens least squares */
static int gmres(size_t m, matop_fn f, void* ctx, const cx* b, cx* x, double tol, int krylovdim,
                 int maxiter, int* matvecs) {
    cx* V = cx_alloc(m * (size_t)(krylovdim + 1));
    cx* H = cx_alloc((size_t)(krylovdim + 1) * krylovdim);
    cx* g = cx_alloc((size_t)krylovdim + 1);
    double* cs = (double*)malloc(sizeof(double) * (size_t)krylovdim);
    cx* sn = cx_alloc((size_t)krylovdim);
    cx* w = cx_alloc(m);
    cx* yk = cx_alloc((size_t)krylovdim);
    memset(x, 0, sizeof(cx) * m);
    int nmv = 0, converged = 0;
    for (int outer = 0; outer < maxiter && !converged; ++outer) {
        f(ctx, x, w); nmv++;
        for (size_t i = 0; i < m; ++i) w[i] = b[i] - w[i];
        double beta = fro_norm(m, w);
        if (beta <= tol) { converged = 1; break; }
        for (size_t i = 0; i < m; ++i) V[i] = w[i] / beta;
        memset(g, 0, sizeof(cx) * (size_t)(krylovdim + 1));
        g[0] = beta;
        int k;
        for (k = 0; k < krylovdim; ++k) {
            f(ctx, V + (size_t)k * m, w); nmv++;
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 0; i <= k; ++i) {
                    cx hij = 0;
                    const cx* vi = V + (size_t)i * m;
                    for (size_t l = 0; l < m; ++l) hij += conj(vi[l]) * w[l];
                    for (size_t l = 0; l < m; ++l) w[l] -= hij * vi[l];
                    H[(size_t)i * krylovdim + k] += hij;
                }
            }
            double hn = fro_norm(m, w);
            H[(size_t)(k + 1) * krylovdim + k] = hn;
            if (hn > 0) for (size_t l = 0; l < m; ++l) V[(size_t)(k + 1) * m + l] = w[l] / hn;
            for (int i = 0; i < k; ++i) {
                cx h1 = H[(size_t)i * krylovdim + k], h2 = H[(size_t)(i + 1) * krylovdim + k];
                H[(size_t)i * krylovdim + k] = cs[i] * h1 + sn[i] * h2;
                H[(size_t)(i + 1) * krylovdim + k] = -conj(sn[i]) * h1 + cs[i] * h2;
            }
            givens(H[(size_t)k * krylovdim + k], H[(size_t)(k + 1) * krylovdim + k], &cs[k], &sn[k]);
            cx h1 = H[(size_t)k * krylovdim + k], h2 = H[(size_t)(k + 1) * krylovdim + k];
            H[(size_t)k * krylovdim + k] = cs[k] * h1 + sn[k] * h2;
            H[(size_t)(k + 1) * krylovdim + k] = 0;
            cx g1 = g[k], g2 = g[k + 1];
            g[k] = cs[k] * g1 + sn[k] * g2;
            g[k + 1] = -conj(sn[k]) * g1 + cs[k] * g2;
            if (cabs(g[k + 1]) <= tol || hn == 0) { ++k; converged = cabs(g[k]) <= tol || hn == 0; break; }
        }
        int kk = k > krylovdim ? krylovdim : k;
        for (int i = kk - 1; i >= 0; --i) {
            cx s = g[i];
            for (int l = i + 1; l < kk; ++l) s -= H[(size_t)i * krylovdim + l] * yk[l];
            yk[i] = s / H[(size_t)i * krylovdim + i];
        }
        for (int i = 0; i < kk; ++i)
            for (size_t l = 0; l < m; ++l) x[l] += yk[i] * V[(size_t)i * m + l];
        memset(H, 0, sizeof(cx) * (size_t)(krylovdim + 1) * krylovdim);
        converged = 0;
        f(ctx, x, w); nmv++;
        for (size_t i = 0; i < m; ++i) w[i] = b[i] - w[i];
        if (fro_norm(m, w) <= tol) converged = 1;
    }
    if (matvecs) *matvecs = nmv;
    free(V); free(H); free(g); free(cs); free(sn); free(w); free(yk);
    return converged ? OR_OK : OR_SOLVER_FAILURE;
}

/* method: 0 = iterative (GMRES + no-jump preconditioner), 1 = dense */
int or_boundary(int n, const cx* q, const cx* r, const cx* rho, const cx* a, double tol, int method,
                cx* h) {
    size_t m = (size_t)n * n;
    cx* b = cx_alloc(m);
    cx t = trace_prod(n, rho, a);
    memcpy(b, a, sizeof(cx) * m);
    for (int i = 0; i < n; ++i) b[(size_t)i * n + i] -= t;
    if (n == 1) { memset(h, 0, sizeof(cx) * m); free(b); return OR_OK; }
    bnd_ctx_t ctx = {n, q, r, rho, NULL, NULL, 0, cx_alloc(m)};
    int st;
    double nb = fro_norm(m, b);
    if (method == 1) {
        st = dense_operator_solve(n, bnd_op, &ctx, b, h);
    } else {
        nojump_t p;
        st = nojump_init(&p, n, q, r);
        if (!st) {
            cx* id = cx_alloc(m);
            cx* z = cx_alloc(m);
            for (int i = 0; i < n; ++i) id[(size_t)i * n + i] = 1;
            nojump_apply(&p, id, z, 1);
            ctx.p = &p; ctx.z = z;
            ctx.den = 1 + trace_prod(n, rho, z);
            cx* y = cx_alloc(m);
            int kd = (int)(m < 40 ? m : 40);
            st = gmres(m, bnd_op_pre, &ctx, b, y, tol * fmax(1.0, nb), kd, 200, NULL);
            bnd_pre(&ctx, y, h);
            free(y); free(id); free(z);
            nojump_free(&p);
        }
    }
    if (!st) {
        cx* res = cx_alloc(m);
        boundary_op(n, q, r, rho, h, res);
        for (size_t i = 0; i < m; ++i) res[i] -= b[i];
        if (!(fro_norm(m, res) <= 20 * tol * fmax(1.0, nb))) st = OR_SOLVER_FAILURE;
        free(res);
    }
    free(ctx.tmp); free(b);
    return st;
}

/* G = [(rho + reg I)^{-1} covector]^dagger  (src/observables.jl:210-211) */
static int metric_solve(int n, const cx* rho, double reg, const cx* cov, cx* grad) {
    size_t m = (size_t)n * n;
    cx* A = cx_alloc(m);
    cx* B = cx_alloc(m);
    memcpy(A, rho, sizeof(cx) * m);
    for (int i = 0; i < n; ++i) A[(size_t)i * n + i] += reg;
    memcpy(B, cov, sizeof(cx) * m);
    int st = lu_solve(n, A, n, B);
    adjoint_of(n, B, grad);
    free(A); free(B);
    return st;
}

/* ------------------------------------------------------------------------------------------ */
/* stationary state (src/transfer.jl:33-47,74-115)                                            */
/* ------------------------------------------------------------------------------------------ */

static void lindblad_void(void* c, const cx* x, cx* y) {
    bnd_ctx_t* b = (bnd_ctx_t*)c;
    lindblad(b->n, b->q, b->r, x, y, 0);
}

typedef struct { int n; const cx *q, *r; } lctx_t;

static void dense_rho_op(void* c, const cx* x, cx* y) {
    /* (T + v v^dagger / D) x with v = vec(I): L(x) + tr(x)/D * I */
    lctx_t* l = (lctx_t*)c;
    int n = l->n;
    lindblad(n, l->q, l->r, x, y, 0);
    cx t = trace_of(n, x) / n;
    for (int i = 0; i < n; ++i) y[(size_t)i * n + i] += t;
}

/* checked_density (src/transfer.jl:33-47); rho normalised in place */
static int checked_density(int n, const cx* q, const cx* r, cx* rho, double tol) {
    size_t m = (size_t)n * n;
    for (size_t i = 0; i < m; ++i) if (!isfinite(creal(rho[i])) || !isfinite(cimag(rho[i]))) return OR_SOLVER_FAILURE;
    cx z = trace_of(n, rho);
    if (!(cabs(z) > 2.220446049250313e-16 * fro_norm(m, rho))) return OR_SOLVER_FAILURE;
    for (size_t i = 0; i < m; ++i) rho[i] /= z;
    double herm = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            cx d = rho[(size_t)i * n + k] - conj(rho[(size_t)k * n + i]);
            herm += creal(d * conj(d));
        }
    if (!(sqrt(herm) <= 100 * tol)) return OR_SOLVER_FAILURE;
    cx* x = cx_alloc(m);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) x[(size_t)i * n + k] = (rho[(size_t)i * n + k] + conj(rho[(size_t)k * n + i])) / 2;
    memcpy(rho, x, sizeof(cx) * m);
    cx* l = cx_alloc(m);
    lindblad(n, q, r, rho, l, 0);
    double res = fro_norm(m, l);
    double nq = opnorm2(n, q), nr = opnorm2(n, r);
    double scale = fmax(1.0, 2 * nq + nr * nr);
    int st = res <= tol * scale ? OR_OK : OR_SOLVER_FAILURE;
    /* positivity: smallest eigenvalue of the Hermitian x from its Schur form */
    cx* Z = cx_alloc(m);
    if (!st && !or_schur(n, x, Z)) {
        double mn = INFINITY;
        for (int i = 0; i < n; ++i) mn = fmin(mn, creal(x[(size_t)i * n + i]));
        if (mn < -10 * tol) st = OR_SOLVER_FAILURE;
    }
    free(Z); free(x); free(l);
    return st;
}

/* method 0: CPTP Arnoldi (restarted at the Ritz vector of largest real part); 1: dense */
int or_steady_state(int n, const cx* q, const cx* r, double tol, int method, cx* rho, int* iterations) {
    size_t m = (size_t)n * n;
    if (n == 1) { rho[0] = 1; if (iterations) *iterations = 0; return OR_OK; }
    int st;
    if (method == 1) {
        lctx_t l = {n, q, r};
        cx* b = cx_alloc(m);
        for (int i = 0; i < n; ++i) b[(size_t)i * n + i] = 1.0 / n;
        st = dense_operator_solve(n, dense_rho_op, &l, b, rho);
        free(b);
        if (iterations) *iterations = 1;
    } else {
        nojump_t p;
        st = nojump_init(&p, n, q, r);
        if (st) return st;
        int kd = (int)(m < 40 ? m : 40);
        cx* V = cx_alloc(m * (size_t)(kd + 1));
        cx* H = cx_alloc((size_t)(kd + 1) * kd);
        cx* Hs = cx_alloc((size_t)kd * kd);
        cx* Z = cx_alloc((size_t)kd * kd);
        cx* w = cx_alloc(m);
        cx* t1 = cx_alloc(m);
        cx* t2 = cx_alloc(m);
        cx* xi = cx_alloc(m);
        cx* ev = cx_alloc((size_t)kd);
        for (int i = 0; i < n; ++i) xi[(size_t)i * n + i] = 1.0 / n;
        int ops = 0, converged = 0;
        st = OR_SOLVER_FAILURE;
        for (int outer = 0; outer < 200 && !converged; ++outer) {
            double nx = fro_norm(m, xi);
            for (size_t i = 0; i < m; ++i) V[i] = xi[i] / nx;
            memset(H, 0, sizeof(cx) * (size_t)(kd + 1) * kd);
            int k;
            for (k = 0; k < kd; ++k) {
                /* Phi(X) = R P(X) R^dagger + shift P(X) */
                nojump_apply(&p, V + (size_t)k * m, t1, 0);
                mm(n, 1, r, 0, t1, 0, 0, t2);
                mm(n, 1, t2, 0, r, 1, 0, w);
                for (size_t i = 0; i < m; ++i) w[i] += p.shift * t1[i];
                ops++;
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = 0; i <= k; ++i) {
                        cx hij = 0;
                        for (size_t l = 0; l < m; ++l) hij += conj(V[(size_t)i * m + l]) * w[l];
                        for (size_t l = 0; l < m; ++l) w[l] -= hij * V[(size_t)i * m + l];
                        H[(size_t)i * kd + k] += hij;
                    }
                double hn = fro_norm(m, w);
                H[(size_t)(k + 1) * kd + k] = hn;
                if (hn < 1e-300) { ++k; break; }
                for (size_t l = 0; l < m; ++l) V[(size_t)(k + 1) * m + l] = w[l] / hn;
            }
            int kk = k;
            /* Ritz pairs of the kk x kk Hessenberg */
            for (int i = 0; i < kk; ++i)
                for (int j = 0; j < kk; ++j) Hs[(size_t)i * kk + j] = H[(size_t)i * kd + j];
            if (or_schur(kk, Hs, Z)) break;
            int best = 0;
            for (int i = 1; i < kk; ++i)
                if (creal(Hs[(size_t)i * kk + i]) > creal(Hs[(size_t)best * kk + best])) best = i;
            cx lam = Hs[(size_t)best * kk + best];
            /* eigenvector of the triangular factor for index best, then rotate by Z */
            cx* s = cx_alloc((size_t)kk);
            s[best] = 1;
            for (int i = best - 1; i >= 0; --i) {
                cx acc = 0;
                for (int l = i + 1; l <= best; ++l) acc += Hs[(size_t)i * kk + l] * s[l];
                cx den = lam - Hs[(size_t)i * kk + i];
                if (cabs(den) < 1e-300) den = 1e-300;
                s[i] = acc / den;
            }
            for (int i = 0; i < kk; ++i) {
                cx acc = 0;
                for (int l = 0; l <= best; ++l) acc += Z[(size_t)i * kk + l] * s[l];
                ev[i] = acc;
            }
            free(s);
            double evn = 0;
            for (int i = 0; i < kk; ++i) evn += creal(ev[i] * conj(ev[i]));
            evn = sqrt(evn);
            for (int i = 0; i < kk; ++i) ev[i] /= evn;
            memset(xi, 0, sizeof(cx) * m);
            for (int i = 0; i < kk; ++i)
                for (size_t l = 0; l < m; ++l) xi[l] += ev[i] * V[(size_t)i * m + l];
            double resid = kk < kd ? 0 : cabs(H[(size_t)kk * kd + kk - 1] * ev[kk - 1]);
            if (resid <= tol / 10 || kk < kd) {
                converged = 1;
                if (cabs(lam - 1) <= 100 * tol) st = OR_OK;
            }
        }
        if (!st) nojump_apply(&p, xi, rho, 0);
        if (iterations) *iterations = ops;
        free(V); free(H); free(Hs); free(Z); free(w); free(t1); free(t2); free(xi); free(ev);
        nojump_free(&p);
        if (st) return st;
    }
    if (st) return st;
    return checked_density(n, q, r, rho, tol);
}

/* ------------------------------------------------------------------------------------------ */
/* trajectories and fg_phi4                                                                   */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int n_quad;
    double max_point, atol, rtol, initial_step;
    long max_steps;
} or_integrator_t;

/* Integrate one system; samples has (2*n_quad+1) * len entries (nodes then endpoint).
   Returns status; stats[0..4] = accepted, rejected, attempts, dense_steps, rhs_calls. */
static int run_system(system_t* sys, const or_integrator_t* ip, const double* save_points, int npts,
                      cx* samples, long* stats) {
    int n = sys->ctx.n;
    size_t m = (size_t)n * n;
    int ch = sys->ctx.spec.channels;
    sysrhs_t w = {sys, cx_alloc(m * ch), cx_alloc(m * ch), 0};
    traj_stats_t st;
    int status = integrate(&w, save_points, npts, ip->atol, ip->rtol, ip->initial_step, ip->max_steps,
                           samples, &st);
    if (stats) {
        stats[0] = st.accepted; stats[1] = st.rejected; stats[2] = st.attempts;
        stats[3] = st.dense_steps; stats[4] = st.rhs_calls;
    }
    free(w.full); free(w.dfull);
    return status;
}

size_t or_system_len(int n, int system) {
    system_t s;
    system_init(&s, n, NULL, NULL, NULL, NULL, system / 2, system % 2);
    return s.len;
}

/* Public: one reduced trajectory (for parity of samples). */
int or_trajectory(int n, const cx* q, const cx* r, const cx* rho, int system, const or_integrator_t* ip,
                  cx* samples, long* stats) {
    size_t m = (size_t)n * n;
    cx* comm = cx_alloc(m);
    mm(n, 1, q, 0, r, 0, 0, comm);
    mm(n, -1, r, 0, q, 0, 1, comm);
    int nn = 2 * ip->n_quad;
    double* pts = (double*)malloc(sizeof(double) * (size_t)(nn + 1));
    double* wts = (double*)malloc(sizeof(double) * (size_t)nn);
    or_quadrature(ip->n_quad, ip->max_point, pts, wts);
    pts[nn] = sqrt(ip->max_point);
    system_t s;
    system_init(&s, n, q, r, comm, rho, system / 2, system % 2);
    int st = run_system(&s, ip, pts, nn + 1, samples, stats);
    free(pts); free(wts); free(comm);
    return st;
}

/* Public: literal full-channel RHS of one spec (for tests of the reduced RHS). */
void or_rhs_full(int n, const cx* q, const cx* r, int term, int adj, double t, const cx* u, cx* du) {
    size_t m = (size_t)n * n;
    cx* comm = cx_alloc(m);
    mm(n, 1, q, 0, r, 0, 0, comm);
    mm(n, -1, r, 0, q, 0, 1, comm);
    rhs_ctx_t c = {n, q, r, term == 1 ? comm : r, term == 2 ? monomial_spec(4) : kinetic_spec(term == 1), adj};
    rhs_full(&c, t, u, du);
    free(comm);
}

/*
 * fg_phi4 (src/observables.jl:226-257): terms (1, kinetic), (1, kinetic commutator), (g, phi^4).
 * Outputs: energy, gradient (D x D), raw covector and environment (before finish_gradient),
 * per-term values, per-system stats (6 x 5 longs).  flags bit0: skip the boundary term;
 * bit1: dense boundary solve.  Returns a status code.
 */
int or_kinetic_resolvent(int n, const cx* r, const cx* q, const cx* rho, int minnodes, int maxnodes, double atol,
                         double rtol, double metric_reg, double boundary_tol, int gradient, double* value, cx* cov,
                         cx* env, int* nodes_used);

static int fg_phi4_core(int n, const cx* r, const cx* q, const cx* rho, double g, const or_integrator_t* ip,
                        double metric_reg, double boundary_tol, int flags, int nthreads,
                        double* energy, cx* grad, cx* raw_cov, cx* raw_env, double* term_values, long* stats,
                        cx* grad_r, cx* grad_k) {
    size_t m = (size_t)n * n;
    if (n < 1 || ip->n_quad < 8 || !(ip->max_point > 0) || !(ip->atol > 0) || !(ip->rtol > 0))
        return OR_BAD_ARGUMENT;
    /* flags bit2: kinetic terms on the infinite line (kinetic_backend=:resolvent, default options) */
    const int kin_res = (flags & 4) != 0;
    if (kin_res && grad_r) return OR_BAD_ARGUMENT;
    cx* comm = cx_alloc(m);
    mm(n, 1, q, 0, r, 0, 0, comm);
    mm(n, -1, r, 0, q, 0, 1, comm);
    int nn = 2 * ip->n_quad;
    double* pts = (double*)malloc(sizeof(double) * (size_t)(nn + 1));
    double* wts = (double*)malloc(sizeof(double) * (size_t)nn);
    or_quadrature(ip->n_quad, ip->max_point, pts, wts);
    pts[nn] = sqrt(ip->max_point);
    system_t sys[6];
    cx* samples[6];
    int status[6];
    for (int s = 0; s < 6; ++s) {
        system_init(&sys[s], n, q, r, comm, rho, s / 2, s % 2);
        samples[s] = cx_alloc(sys[s].len * (size_t)(nn + 1));
    }
    long st_local[30];
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 1)
    for (int s = 0; s < 6; ++s) {
        if (kin_res && s < 4) {
            status[s] = OR_OK;
            memset(st_local + 5 * s, 0, 5 * sizeof(long));
        } else {
            status[s] = run_system(&sys[s], ip, pts, nn + 1, samples[s], st_local + 5 * s);
        }
    }
    if (stats) memcpy(stats, st_local, sizeof st_local);
    int st = OR_OK;
    for (int s = 0; s < 6 && !st; ++s) st = status[s];
    double coeff[3] = {1.0, 1.0, g};
    double value = 0;
    cx* cov = cx_alloc(m);
    cx* env = cx_alloc(m);
    cx* c1 = cx_alloc(m);
    cx* e1 = cx_alloc(m);
    cx *xr = NULL, *xq = NULL, *xr1 = NULL, *xq1 = NULL;
    if (grad_r) { xr = cx_alloc(m); xq = cx_alloc(m); xr1 = cx_alloc(m); xq1 = cx_alloc(m); }
    if (!st && kin_res) {   /* fixed reduction order: the kinetic job first (src/observables.jl:236-246) */
        double kv;
        st = or_kinetic_resolvent(n, r, q, rho, 32, 512, 1e-11, 1e-10, metric_reg, boundary_tol, 1, &kv, c1, e1, NULL);
        if (!st) {
            value += kv;
            for (size_t i = 0; i < m; ++i) { cov[i] += c1[i]; env[i] += e1[i]; }
            if (term_values) { term_values[0] = kv; term_values[1] = 0; }
        }
    }
    if (!st) {
        for (int term = kin_res ? 2 : 0; term < 3; ++term) {
            double v;
            pullback(&sys[2 * term], &sys[2 * term + 1], samples[2 * term], samples[2 * term + 1],
                     pts, wts, nn, c1, e1, &v, xr1, xq1);
            if (term_values) term_values[term] = v;
            value += coeff[term] * v;
            for (size_t i = 0; i < m; ++i) { cov[i] += coeff[term] * c1[i]; env[i] += coeff[term] * e1[i]; }
            if (xr)
                for (size_t i = 0; i < m; ++i) { xr[i] += coeff[term] * xr1[i]; xq[i] += coeff[term] * xq1[i]; }
        }
        if (raw_cov) memcpy(raw_cov, cov, sizeof(cx) * m);
        if (raw_env) memcpy(raw_env, env, sizeof(cx) * m);
        *energy = value;
        if (!(flags & 1)) {
            cx* h = cx_alloc(m);
            cx* tmp = cx_alloc(m);
            st = or_boundary(n, q, r, rho, env, boundary_tol, (flags & 2) ? 1 : 0, h);
            if (!st) {
                addtriple(n, cov, rho, 0, r, 1, h, 0, 1, tmp);
                addtriple(n, cov, rho, 0, h, 0, r, 1, -1, tmp);
                if (xr) {   /* rho0 responds to dQ rho0 + rho0 dQ^+ + dR rho0 R^+ + R rho0 dR^+ */
                    addtriple(n, xr, rho, 0, r, 1, h, 0, 1, tmp);
                    mm(n, 1, rho, 0, h, 0, 1, xq);
                }
            }
            free(h); free(tmp);
        }
        if (!st && grad) st = metric_solve(n, rho, metric_reg, cov, grad);
        if (!st && grad_r) {
            /* dQ = -i dK - (dR^+ R + R^+ dR)/2:  grad_R = 2 (xr - herm(xq) R^+)^+,
               grad_K = 2 herm(-i xq) = -i (xq - xq^+) */
            cx* t = cx_alloc(m);
            cx* hq = cx_alloc(m);
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) hq[(size_t)i * n + k] = (xq[(size_t)i * n + k] + conj(xq[(size_t)k * n + i])) / 2;
            memcpy(t, xr, sizeof(cx) * m);
            mm(n, -1, hq, 0, r, 1, 1, t);
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    grad_r[(size_t)i * n + k] = 2 * conj(t[(size_t)k * n + i]);
                    grad_k[(size_t)i * n + k] = -I * (xq[(size_t)i * n + k] - conj(xq[(size_t)k * n + i]));
                }
            free(t); free(hq);
        }
    }
    for (int s = 0; s < 6; ++s) free(samples[s]);
    free(cov); free(env); free(c1); free(e1); free(pts); free(wts); free(comm);
    free(xr); free(xq); free(xr1); free(xq1);
    return st;
}

int or_fg_phi4(int n, const cx* r, const cx* q, const cx* rho, double g, const or_integrator_t* ip,
               double metric_reg, double boundary_tol, int flags, int nthreads,
               double* energy, cx* grad, cx* raw_cov, cx* raw_env, double* term_values, long* stats) {
    return fg_phi4_core(n, r, q, rho, g, ip, metric_reg, boundary_tol, flags, nthreads, energy, grad, raw_cov,
                        raw_env, term_values, stats, NULL, NULL);
}

/*
 * Energy and Euclidean gradient in (R, K) for Q = -iK - R^+R/2 with rho0 fixed by (R, K):
 * dE = Re sum conj(grad_r) dR + Re sum conj(grad_k) dK for Hermitian dK.  The reference has only
 * the Riemannian gradient (src/observables.jl:161-212); this keeps the dR and dQ coefficients of
 * the same pullback apart.
 */
int or_fg_phi4_rk(int n, const cx* r, const cx* q, const cx* rho, double g, const or_integrator_t* ip,
                  double boundary_tol, int flags, int nthreads, double* energy, cx* grad_r, cx* grad_k) {
    if (!grad_r || !grad_k) return OR_BAD_ARGUMENT;
    return fg_phi4_core(n, r, q, rho, g, ip, 0.0, boundary_tol, flags, nthreads, energy, NULL, NULL, NULL, NULL,
                        NULL, grad_r, grad_k);
}

/* ------------------------------------------------------------------------------------------ */
/* infinite-line kinetic terms by the angular resolvent (src/kinetic.jl:1-124), dense Schur     */
/* ------------------------------------------------------------------------------------------ */

/* transfer matrix in row-major vec: vec(L(X)) = T vec(X)  (src/transfer.jl:21-26 up to the vec order) */
static void transfer_matrix(int n, const cx* q, const cx* r, cx* T) {
    size_t m = (size_t)n * n;
    cx* e = cx_alloc(m);
    cx* col = cx_alloc(m);
    for (size_t j = 0; j < m; ++j) {
        memset(e, 0, sizeof(cx) * m);
        e[j] = 1;
        lindblad(n, q, r, e, col, 0);
        for (size_t i = 0; i < m; ++i) T[i * m + j] = col[i];
    }
    free(e); free(col);
}

/* Gauss-Legendre nodes on (0, pi/2) with weights summing to one (the factor 2/pi included),
   src/kinetic.jl:19-23 (there by Golub-Welsch) */
static void kinetic_grid(int nodes, double* theta, double* w) {
    gauss_legendre(nodes, theta, w);
    for (int i = 0; i < nodes; ++i) {
        theta[i] = M_PI / 4 * (theta[i] + 1);
        w[i] /= 2;
    }
}

/* One kinetic spec: (2/pi) int_0^{pi/2} tr[op (1 - cos(theta) L)^{-1} (rho op^+)] dtheta, with its
   covector and environment (src/kinetic.jl:25-60).  U, Z: Schur factors of T (T = Z U Z^+). */
static void kinetic_components(int n, const cx* q, const cx* r, const cx* rho, int commutator, const cx* U,
                               const cx* Z, int nodes, const double* theta, const double* wts, int gradient,
                               double* value, cx* cov, cx* env) {
    size_t m = (size_t)n * n;
    cx *op = cx_alloc(m), *opd = cx_alloc(m), *rhoopd = cx_alloc(m), *b = cx_alloc(m), *c = cx_alloc(m);
    cx *xx = cx_alloc(m), *yy = cx_alloc(m), *x = cx_alloc(m), *A = cx_alloc(m), *a = cx_alloc(m);
    cx *M = cx_alloc(m), *N = cx_alloc(m), *u = cx_alloc(m), *v = cx_alloc(m), *E = cx_alloc(m), *tmp = cx_alloc(m);
    if (commutator) { mm(n, 1, q, 0, r, 0, 0, op); mm(n, -1, r, 0, q, 0, 1, op); }
    else memcpy(op, r, sizeof(cx) * m);
    adjoint_of(n, op, opd);
    mm(n, 1, rho, 0, opd, 0, 0, rhoopd);
    for (size_t i = 0; i < m; ++i) {      /* b = Z^+ vec(rho op^+), c = Z^+ vec(op^+) */
        cx sb = 0, sc = 0;
        for (size_t k = 0; k < m; ++k) { sb += conj(Z[k * m + i]) * rhoopd[k]; sc += conj(Z[k * m + i]) * opd[k]; }
        b[i] = sb; c[i] = sc;
    }
    double val = 0;
    for (int t = 0; t < nodes; ++t) {
        const double z = cos(theta[t]), w = wts[t];
        /* (I - z U) xx = b, upper triangular */
        for (long i = (long)m - 1; i >= 0; --i) {
            cx sacc = b[i];
            for (size_t k = (size_t)i + 1; k < m; ++k) sacc += z * U[(size_t)i * m + k] * xx[k];
            xx[i] = sacc / (1.0 - z * U[(size_t)i * m + i]);
        }
        cx d = 0;
        for (size_t i = 0; i < m; ++i) d += conj(c[i]) * xx[i];
        val += w * creal(d);
        if (!gradient) continue;
        /* (I - z U)^+ yy = c, lower triangular */
        for (size_t i = 0; i < m; ++i) {
            cx sacc = c[i];
            for (size_t k = 0; k < i; ++k) sacc += z * conj(U[k * m + i]) * yy[k];
            yy[i] = sacc / (1.0 - z * conj(U[i * m + i]));
        }
        for (size_t i = 0; i < m; ++i) {
            cx sx = 0, sa = 0;
            for (size_t k = 0; k < m; ++k) { sx += Z[i * m + k] * xx[k]; sa += Z[i * m + k] * yy[k]; }
            x[i] = sx; A[i] = sa;
        }
        adjoint_of(n, A, a);
        addtriple(n, M, x, 0, r, 1, a, 0, w * z, tmp);
        addtriple(n, M, x, 0, a, 0, r, 1, -w * z, tmp);
        addtriple(n, N, a, 0, r, 0, x, 0, w * z, tmp);
        addtriple(n, N, r, 0, a, 0, x, 0, -w * z, tmp);
        for (size_t i = 0; i < m; ++i) u[i] += w * x[i];
        mm(n, w, a, 0, rho, 0, 1, v);
        mm(n, w, opd, 0, a, 0, 1, E);
    }
    *value = val;
    if (gradient) {
        if (commutator) {
            mm(n, 1, u, 0, q, 0, 1, M); mm(n, -1, q, 0, u, 0, 1, M);
            addtriple(n, M, r, 0, u, 0, r, 1, -1, tmp);
            addtriple(n, M, u, 0, r, 0, r, 1, 1, tmp);
            mm(n, 1, q, 1, v, 0, 1, N); mm(n, -1, v, 0, q, 1, 1, N);
            addtriple(n, N, r, 0, v, 0, r, 1, -1, tmp);
            addtriple(n, N, r, 0, r, 1, v, 0, 1, tmp);
        } else {
            for (size_t i = 0; i < m; ++i) { M[i] += u[i]; N[i] += v[i]; }
        }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                size_t ik = (size_t)i * n + k, ki = (size_t)k * n + i;
                cov[ik] = (M[ik] + conj(N[ki])) / 2;
                env[ik] = (E[ik] + conj(E[ki])) / 2;
            }
    }
    free(op); free(opd); free(rhoopd); free(b); free(c); free(xx); free(yy); free(x); free(A); free(a);
    free(M); free(N); free(u); free(v); free(E); free(tmp);
}

/* kinetic1 + kinetic2 at a fixed node count (src/kinetic.jl:62-75 with weights (1, 1)) */
static void kinetic_sum(int n, const cx* q, const cx* r, const cx* rho, const cx* U, const cx* Z, int nodes,
                        int gradient, double* value, cx* cov, cx* env) {
    size_t m = (size_t)n * n;
    double* th = (double*)malloc(sizeof(double) * (size_t)nodes);
    double* w = (double*)malloc(sizeof(double) * (size_t)nodes);
    kinetic_grid(nodes, th, w);
    cx *c1 = cx_alloc(m), *e1 = cx_alloc(m);
    *value = 0;
    if (gradient) { memset(cov, 0, sizeof(cx) * m); memset(env, 0, sizeof(cx) * m); }
    for (int comm = 0; comm < 2; ++comm) {
        double v;
        kinetic_components(n, q, r, rho, comm, U, Z, nodes, th, w, gradient, &v, c1, e1);
        *value += v;
        if (gradient)
            for (size_t i = 0; i < m; ++i) { cov[i] += c1[i]; env[i] += e1[i]; }
    }
    free(th); free(w); free(c1); free(e1);
}

/* finish_gradient (src/observables.jl:201-212) of a covector/environment pair */
static int finish_gradient(int n, const cx* q, const cx* r, const cx* rho, const cx* cov, const cx* env,
                           double metric_reg, double boundary_tol, cx* grad) {
    size_t m = (size_t)n * n;
    cx *c = cx_alloc(m), *h = cx_alloc(m), *tmp = cx_alloc(m);
    memcpy(c, cov, sizeof(cx) * m);
    int st = or_boundary(n, q, r, rho, env, boundary_tol, 1, h);
    if (!st) {
        addtriple(n, c, rho, 0, r, 1, h, 0, 1, tmp);
        addtriple(n, c, rho, 0, h, 0, r, 1, -1, tmp);
        st = metric_solve(n, rho, metric_reg, c, grad);
    }
    free(c); free(h); free(tmp);
    return st;
}

/*
 * kinetic_resolvent (src/kinetic.jl:77-124) for the two kinetic terms: node doubling from minnodes
 * until the energy and the finished gradient change by at most atol + rtol |.|.  Returns the raw
 * value, covector and environment of the accepted level and the node count in *nodes_used.
 */
int or_kinetic_resolvent(int n, const cx* r, const cx* q, const cx* rho, int minnodes, int maxnodes, double atol,
                         double rtol, double metric_reg, double boundary_tol, int gradient, double* value, cx* cov,
                         cx* env, int* nodes_used) {
    if (n < 1 || minnodes < 2 || maxnodes < 2 * minnodes || !(atol > 0) || !(rtol >= 0)) return OR_BAD_ARGUMENT;
    size_t m = (size_t)n * n;
    cx *U = cx_alloc(m * m), *Z = cx_alloc(m * m);
    transfer_matrix(n, q, r, U);
    int st = or_schur((int)m, U, Z);
    if (st) { free(U); free(Z); return st; }
    cx *pc = cx_alloc(m), *pe = cx_alloc(m), *pg = cx_alloc(m), *g = cx_alloc(m);
    int nodes = minnodes;
    double pv;
    kinetic_sum(n, q, r, rho, U, Z, nodes, gradient, &pv, pc, pe);
    if (gradient) st = finish_gradient(n, q, r, rho, pc, pe, metric_reg, boundary_tol, pg);
    int done = 0;
    while (!st && !done && 2 * nodes <= maxnodes) {
        nodes *= 2;
        double v;
        kinetic_sum(n, q, r, rho, U, Z, nodes, gradient, &v, cov, env);
        if (!isfinite(v)) { st = OR_NONFINITE; break; }
        int energy_pass = fabs(v - pv) <= atol + rtol * fabs(v);
        if (gradient) {
            st = finish_gradient(n, q, r, rho, cov, env, metric_reg, boundary_tol, g);
            if (st) break;
            double dg = 0, gn = 0;
            for (size_t i = 0; i < m; ++i) { dg += cabs(g[i] - pg[i]) * cabs(g[i] - pg[i]); gn += cabs(g[i]) * cabs(g[i]); }
            dg = sqrt(dg); gn = sqrt(gn);
            if (energy_pass && dg <= atol + rtol * gn) done = 1;
            memcpy(pg, g, sizeof(cx) * m);
        } else if (energy_pass) {
            done = 1;
        }
        pv = v;
        *value = v;
    }
    if (!st && !done) st = OR_SOLVER_FAILURE;
    if (nodes_used) *nodes_used = nodes;
    free(U); free(Z); free(pc); free(pe); free(pg); free(g);
    return st;
}

/* Exposed helpers for tests */
int or_nojump(int n, const cx* q, const cx* r, const cx* y, int adjoint, cx* out) {
    nojump_t p;
    int st = nojump_init(&p, n, q, r);
    if (!st) st = nojump_apply(&p, y, out, adjoint);
    nojump_free(&p);
    return st;
}

int or_metric_solve(int n, const cx* rho, double reg, const cx* cov, cx* grad) {
    return metric_solve(n, rho, reg, cov, grad);
}

void or_lindblad(int n, const cx* q, const cx* r, const cx* x, cx* out, int adjoint) {
    lindblad(n, q, r, x, out, adjoint);
}
