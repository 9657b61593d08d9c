// Device kernels of the tangent-space excitation action: the momentum-shifted no-jump
// preconditioner, Gram-Schmidt over sets of Arnoldi bases, the channel chain of the :phi^4:
// action ODE, DOP853 stage combinations and error norms, and batched trace products.
#include "device_common.h"
#include "excitation.h"

// Z_ij = -Y_ij / den_ij with den = lam_i + conj(lam_j) - (s_re + i s_im) (forward) or
// conj(lam_i) + lam_j - (s_re + i s_im) (adjoint)
__global__ void rc_sylvester_diag_c(int Dp, int D, int adjoint, double s_re, double s_im, const double* lam,
                                    const double* Y, double* Z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Dp * Dp) return;
    const int r = i / Dp, c = i % Dp;
    cplx* z = reinterpret_cast<cplx*>(Z);
    if (r >= D || c >= D) { z[i] = cmake(0, 0); return; }
    const cplx* l = reinterpret_cast<const cplx*>(lam);
    cplx den = adjoint ? cadd(cconj(l[r]), l[c]) : cadd(l[r], cconj(l[c]));
    den.re -= s_re;
    den.im -= s_im;
    const cplx y = reinterpret_cast<const cplx*>(Y)[i];
    const double m = cabs2(den);
    const cplx q = cmake((y.re * den.re + y.im * den.im) / m, (y.im * den.re - y.re * den.im) / m);
    z[i] = cmake(-q.re, -q.im);
}

// out[b * ldo + k] = <V_{b,k}, W_b> (conjugate-linear in V); grid (nk, nbases)
__global__ void __launch_bounds__(256) rc_mdots(const double* V, long long sb, long long sv, int nk, const double* W,
                                                long long sw, int n, double* out, int ldo) {
    __shared__ double sre[256], sim[256];
    const int k = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
    const cplx* v = reinterpret_cast<const cplx*>(V) + (long long)b * sb + (long long)k * sv;
    const cplx* w = reinterpret_cast<const cplx*>(W) + (long long)b * sw;
    double re = 0, im = 0;
    for (int i = t; i < n; i += 256) {
        const cplx a = v[i], c = w[i];
        re += a.re * c.re + a.im * c.im;
        im += a.re * c.im - a.im * c.re;
    }
    sre[t] = re;
    sim[t] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) { sre[t] += sre[t + s]; sim[t] += sim[t + s]; }
        __syncthreads();
    }
    if (t == 0) {
        out[2 * ((long long)b * ldo + k)] = sre[0];
        out[2 * ((long long)b * ldo + k) + 1] = sim[0];
    }
}

// W_b -= sum_{k < nk} h[b * ldh + k] V_{b,k}; grid (ceil(n / 256), nbases)
__global__ void rc_msubtract(const double* V, long long sb, long long sv, int nk, const double* h, int ldh, int n,
                             double* W, long long sw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (i >= n) return;
    const cplx* v = reinterpret_cast<const cplx*>(V) + (long long)b * sb;
    const cplx* hh = reinterpret_cast<const cplx*>(h) + (long long)b * ldh;
    cplx* w = reinterpret_cast<cplx*>(W) + (long long)b * sw;
    cplx s = w[i];
    for (int k = 0; k < nk; ++k) s = csub(s, cmul(hh[k], v[(long long)k * sv + i]));
    w[i] = s;
}

// out_b = W_b / sqrt(Re nrm2[b * ldn]) (a zero norm writes zeros); grid (ceil(n / 256), nbases)
__global__ void rc_mnormalize(int n, const double* W, long long sw, const double* nrm2, int ldn, double* out,
                              long long so) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (i >= n) return;
    const double q = nrm2[2 * (long long)b * ldn];
    const double s = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    reinterpret_cast<cplx*>(out)[(long long)b * so + i] =
        cscale(s, reinterpret_cast<const cplx*>(W)[(long long)b * sw + i]);
}

// Channel chain of the :phi^4: action: F[t, k] += c_k jt (A + B + C)[t, k - dir] with c_k = k for
// the forward edges (dir = 1, k >= 1) and c_k = k + 1 for the reversed edges (dir = -1,
// k <= nch - 2).  Matrices of trajectory t, channel k at (t nch + k) n.  grid (ceil(n/256), ntraj nch)
__global__ void rc_chain(int n, int nch, int dir, double jt, double* F, const double* A, const double* B,
                         const double* C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int tk = blockIdx.y, k = tk % nch;
    if (i >= n) return;
    const int src = k - dir;
    if (src < 0 || src >= nch) return;
    const double coef = (dir > 0 ? k : k + 1) * jt;
    const long long o = (long long)(tk - dir) * n + i;
    cplx s = reinterpret_cast<const cplx*>(A)[o];
    if (B) s = cadd(s, reinterpret_cast<const cplx*>(B)[o]);
    if (C) s = cadd(s, reinterpret_cast<const cplx*>(C)[o]);
    cplx* f = reinterpret_cast<cplx*>(F) + (long long)tk * n + i;
    *f = cadd(*f, cscale(coef, s));
}

// out = y + h sum_{j < c.nk} c.a[j] K_j  (N complex, K_j at K + j sk)
__global__ void rc_stage(long long N, const double* y, const double* K, long long sk, RcStageCoef c, double h,
                         double* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const cplx* kk = reinterpret_cast<const cplx*>(K);
    cplx s = cmake(0, 0);
    for (int j = 0; j < c.nk; ++j)
        if (c.a[j] != 0.0) s = cadd(s, cscale(c.a[j], kk[(long long)j * sk + i]));
    const cplx yy = reinterpret_cast<const cplx*>(y)[i];
    reinterpret_cast<cplx*>(out)[i] = cadd(yy, cscale(h, s));
}

// DOP853 error sums of one segment of `len` complex elements per workgroup (segment = blockIdx.x):
// out[2 s] = sum |E5.K / sc|^2, out[2 s + 1] = sum |E3.K / sc|^2 with sc = atol + rtol max(|y|, |ynew|)
__global__ void __launch_bounds__(256) rc_step_error(int len, const double* y, const double* ynew, const double* K,
                                                     long long sk, RcStageCoef e5, RcStageCoef e3, double atol,
                                                     double rtol, double* out) {
    __shared__ double s5[256], s3[256];
    const int seg = blockIdx.x, t = threadIdx.x;
    const long long base = (long long)seg * len;
    const cplx* kk = reinterpret_cast<const cplx*>(K);
    const cplx* y0 = reinterpret_cast<const cplx*>(y);
    const cplx* y1 = reinterpret_cast<const cplx*>(ynew);
    double a5 = 0, a3 = 0;
    for (int e = t; e < len; e += 256) {
        const long long i = base + e;
        cplx r5 = cmake(0, 0), r3 = cmake(0, 0);
        for (int j = 0; j < e5.nk; ++j) {
            const cplx kj = kk[(long long)j * sk + i];
            if (e5.a[j] != 0.0) r5 = cadd(r5, cscale(e5.a[j], kj));
            if (e3.a[j] != 0.0) r3 = cadd(r3, cscale(e3.a[j], kj));
        }
        const double sc = atol + rtol * fmax(cabsd(y0[i]), cabsd(y1[i]));
        a5 += cabs2(r5) / (sc * sc);
        a3 += cabs2(r3) / (sc * sc);
    }
    s5[t] = a5;
    s3[t] = a3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) { s5[t] += s5[t + h]; s3[t] += s3[t + h]; }
        __syncthreads();
    }
    if (t == 0) { out[2 * seg] = s5[0]; out[2 * seg + 1] = s3[0]; }
}

// out[b] = tr(A_b B_b) for Dp x Dp matrices; grid (nb)
__global__ void __launch_bounds__(256) rc_trace_prod_batched(int Dp, const double* A, long long sa, const double* B,
                                                             long long sb, double* out) {
    __shared__ double sre[256], sim[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const cplx* a = reinterpret_cast<const cplx*>(A) + (long long)b * sa;
    const cplx* c = reinterpret_cast<const cplx*>(B) + (long long)b * sb;
    double re = 0, im = 0;
    for (int idx = t; idx < Dp * Dp; idx += 256) {
        const int i = idx / Dp, k = idx % Dp;
        const cplx p = cmul(a[i * Dp + k], c[k * Dp + i]);
        re += p.re;
        im += p.im;
    }
    sre[t] = re;
    sim[t] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) { sre[t] += sre[t + s]; sim[t] += sim[t + s]; }
        __syncthreads();
    }
    if (t == 0) { out[2 * b] = sre[0]; out[2 * b + 1] = sim[0]; }
}

// X += (c_re + i c_im) Z elementwise over n complex (c by value)
__global__ void rc_axpy_c(int n, double c_re, double c_im, const double* Z, double* X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx* x = reinterpret_cast<cplx*>(X);
    x[i] = cadd(x[i], cmul(cmake(c_re, c_im), reinterpret_cast<const cplx*>(Z)[i]));
}
