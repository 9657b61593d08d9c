// Host dense complex factorisations for the per-state spectral data (see host_la.h).
#include "host_la.h"

#include <cmath>
#include <limits>

namespace rc {

static inline cd at(const Mat& A, int n, int i, int j) { return A[(size_t)i * n + j]; }

void matmul(int n, const Mat& A, const Mat& B, Mat& C) {
    C.assign((size_t)n * n, cd(0, 0));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            cd a = A[(size_t)i * n + k];
            if (a == cd(0, 0)) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
        }
}

Mat adjoint(int n, const Mat& A) {
    Mat B((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) B[(size_t)i * n + j] = std::conj(A[(size_t)j * n + i]);
    return B;
}

// rotation G = [c s; -conj(s) c] with G [x; y] = [r; 0]
static void rotation(cd x, cd y, double& c, cd& s) {
    double ax = std::abs(x), ay = std::abs(y);
    if (ay == 0) { c = 1; s = 0; return; }
    if (ax == 0) { c = 0; s = std::conj(y) / ay; return; }
    double r = std::hypot(ax, ay);
    c = ax / r;
    s = (x / ax) * std::conj(y) / r;
}

bool schur(int n, Mat& A, Mat& Z) {
    Z.assign((size_t)n * n, cd(0, 0));
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1;
    std::vector<cd> v(n);
    // Householder reduction to upper Hessenberg form
    for (int k = 0; k + 2 < n; ++k) {
        double alpha = 0;
        for (int i = k + 1; i < n; ++i) alpha += std::norm(A[(size_t)i * n + k]);
        alpha = std::sqrt(alpha);
        if (alpha == 0) continue;
        cd x0 = A[(size_t)(k + 1) * n + k];
        cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1, 0);
        std::fill(v.begin(), v.end(), cd(0, 0));
        for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
        v[k + 1] += ph * alpha;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        if (vn == 0) continue;
        const double f = 2.0 / vn;
        for (int j = 0; j < n; ++j) {
            cd s = 0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * A[(size_t)i * n + j];
            s *= f;
            for (int i = k + 1; i < n; ++i) A[(size_t)i * n + j] -= v[i] * s;
        }
        for (int i = 0; i < n; ++i) {
            cd s = 0;
            for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
            s *= f;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= s * std::conj(v[j]);
        }
        for (int i = 0; i < n; ++i) {
            cd s = 0;
            for (int j = k + 1; j < n; ++j) s += Z[(size_t)i * n + j] * v[j];
            s *= f;
            for (int j = k + 1; j < n; ++j) Z[(size_t)i * n + j] -= s * std::conj(v[j]);
        }
        for (int i = k + 2; i < n; ++i) A[(size_t)i * n + k] = 0;
    }
    // single-shift QR with Wilkinson shifts and deflation
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0) {
        int lo = hi;
        while (lo > 0) {
            double s = std::abs(A[(size_t)(lo - 1) * n + lo - 1]) + std::abs(A[(size_t)lo * n + lo]);
            if (s == 0) s = 1;
            if (std::abs(A[(size_t)lo * n + lo - 1]) <= eps * s) { A[(size_t)lo * n + lo - 1] = 0; break; }
            --lo;
        }
        if (lo == hi) { --hi; iter = 0; continue; }
        if (++total > 100 * n + 1000) return false;
        ++iter;
        cd a = A[(size_t)(hi - 1) * n + hi - 1], b = A[(size_t)(hi - 1) * n + hi];
        cd c = A[(size_t)hi * n + hi - 1], d = A[(size_t)hi * n + hi];
        cd tr = a + d, det = a * d - b * c;
        cd disc = std::sqrt(tr * tr / 4.0 - det);
        cd m1 = tr / 2.0 + disc, m2 = tr / 2.0 - disc;
        cd mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        if (iter % 11 == 10) mu = d + std::abs(A[(size_t)hi * n + hi - 1]) * 1.5;
        cd x = A[(size_t)lo * n + lo] - mu, y = A[(size_t)(lo + 1) * n + lo];
        for (int k = lo; k < hi; ++k) {
            if (k > lo) { x = A[(size_t)k * n + k - 1]; y = A[(size_t)(k + 1) * n + k - 1]; }
            double cg; cd sg;
            rotation(x, y, cg, sg);
            for (int j = (k > lo ? k - 1 : lo); j < n; ++j) {
                cd t1 = A[(size_t)k * n + j], t2 = A[(size_t)(k + 1) * n + j];
                A[(size_t)k * n + j] = cg * t1 + sg * t2;
                A[(size_t)(k + 1) * n + j] = -std::conj(sg) * t1 + cg * t2;
            }
            int imax = std::min(k + 2, hi);
            for (int i = 0; i <= imax; ++i) {
                cd t1 = A[(size_t)i * n + k], t2 = A[(size_t)i * n + k + 1];
                A[(size_t)i * n + k] = cg * t1 + std::conj(sg) * t2;
                A[(size_t)i * n + k + 1] = -sg * t1 + cg * t2;
            }
            for (int i = 0; i < n; ++i) {
                cd t1 = Z[(size_t)i * n + k], t2 = Z[(size_t)i * n + k + 1];
                Z[(size_t)i * n + k] = cg * t1 + std::conj(sg) * t2;
                Z[(size_t)i * n + k + 1] = -sg * t1 + cg * t2;
            }
            if (k > lo) A[(size_t)(k + 1) * n + k - 1] = 0;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) A[(size_t)i * n + j] = 0;
    return true;
}

void triangular_eigenvectors(int n, const Mat& T, Mat& W) {
    W.assign((size_t)n * n, cd(0, 0));
    double tnorm = 0;
    for (auto& z : T) tnorm = std::max(tnorm, std::abs(z));
    const double smin = std::max(tnorm * std::numeric_limits<double>::epsilon(), 1e-300);
    for (int k = 0; k < n; ++k) {
        const cd lam = T[(size_t)k * n + k];
        W[(size_t)k * n + k] = 1;
        for (int i = k - 1; i >= 0; --i) {
            cd s = 0;
            for (int l = i + 1; l <= k; ++l) s += T[(size_t)i * n + l] * W[(size_t)l * n + k];
            cd den = lam - T[(size_t)i * n + i];
            if (std::abs(den) < smin) den = smin;
            W[(size_t)i * n + k] = s / den;
        }
        // normalise the column
        double nrm = 0;
        for (int i = 0; i <= k; ++i) nrm += std::norm(W[(size_t)i * n + k]);
        nrm = std::sqrt(nrm);
        for (int i = 0; i <= k; ++i) W[(size_t)i * n + k] /= nrm;
    }
}

bool upper_inverse(int n, const Mat& U, Mat& X) {
    X.assign((size_t)n * n, cd(0, 0));
    for (int j = 0; j < n; ++j) {
        if (U[(size_t)j * n + j] == cd(0, 0)) return false;
        X[(size_t)j * n + j] = 1.0 / U[(size_t)j * n + j];
        for (int i = j - 1; i >= 0; --i) {
            cd s = 0;
            for (int l = i + 1; l <= j; ++l) s += U[(size_t)i * n + l] * X[(size_t)l * n + j];
            X[(size_t)i * n + j] = -s / U[(size_t)i * n + i];
        }
    }
    return true;
}

bool inverse(int n, const Mat& A_in, Mat& X) {
    Mat A = A_in;
    X.assign((size_t)n * n, cd(0, 0));
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(A[(size_t)i * n + k]) > best) { best = std::abs(A[(size_t)i * n + k]); p = i; }
        if (best == 0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
                std::swap(X[(size_t)k * n + j], X[(size_t)p * n + j]);
            }
        const cd inv = 1.0 / A[(size_t)k * n + k];
        for (int j = 0; j < n; ++j) { A[(size_t)k * n + j] *= inv; X[(size_t)k * n + j] *= inv; }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const cd f = A[(size_t)i * n + k];
            if (f == cd(0, 0)) continue;
            for (int j = 0; j < n; ++j) {
                A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
                X[(size_t)i * n + j] -= f * X[(size_t)k * n + j];
            }
        }
    }
    return true;
}

double opnorm2(int n, const Mat& A) {
    std::vector<cd> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * i;
    double lam = 0;
    for (int it = 0; it < 300; ++it) {
        for (int i = 0; i < n; ++i) {
            cd s = 0;
            for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * x[k];
            y[i] = s;
        }
        for (int k = 0; k < n; ++k) {
            cd s = 0;
            for (int i = 0; i < n; ++i) s += std::conj(A[(size_t)i * n + k]) * y[i];
            x[k] = s;
        }
        double nx = 0;
        for (auto& z : x) nx += std::norm(z);
        nx = std::sqrt(nx);
        if (nx == 0) break;
        double prev = lam;
        lam = std::sqrt(nx);
        for (auto& z : x) z /= nx;
        if (it > 10 && std::fabs(lam - prev) <= 1e-14 * lam) break;
    }
    return lam;
}

}  // namespace rc
