// Device kernels of the tangent-space excitation action: the momentum-shifted no-jump
// preconditioner, Gram-Schmidt over sets of Arnoldi bases, the fused right-hand side of the :phi^4:
// action ODEs, DOP853 stage combinations and error norms, and batched trace products.
#include "device_common.h"
#include "excitation.h"
#include "trajectory.h"

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
    // thread t adds elements t, t + 256, ... in increasing order, eight elements' loads in flight
    double re = 0, im = 0;
    for (int i0 = t; i0 < n; i0 += 8 * 256) {
        cplx a[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 256 * u < n) {
                a[u] = v[i0 + 256 * u];
                c[u] = w[i0 + 256 * u];
            }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 256 * u < n) {
                re += a[u].re * c[u].re + a[u].im * c[u].im;
                im += a[u].re * c[u].im - a[u].im * c[u].re;
            }
    }
    tree_sum_256(re, im, sre, sim, out + 2 * ((long long)b * ldo + k));
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
    // eight vectors' loads in flight, subtracted in increasing k
    for (int k0 = 0; k0 < nk; k0 += 8) {
        cplx x[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nk) {
                x[u] = v[(long long)(k0 + u) * sv + i];
                c[u] = hh[k0 + u];
            }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < nk) s = csub(s, cmul(c[u], x[u]));
    }
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

// out = y + h sum_{j < c.nk} c.a[j] K_j  (N complex, K_j at K + j sk); h from hp when non-null
__global__ void rc_stage(long long N, const double* y, const double* K, long long sk, RcStageCoef c, double h,
                         const double* hp, double* out, const int* active) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || (active && !*active)) return;
    if (hp) h = hp[0];
    const cplx* kk = reinterpret_cast<const cplx*>(K);
    cplx s = cmake(0, 0);
    for (int j = 0; j < c.nk; ++j)
        if (c.a[j] != 0.0) s = cadd(s, cscale(c.a[j], kk[(long long)j * sk + i]));
    const cplx yy = reinterpret_cast<const cplx*>(y)[i];
    reinterpret_cast<cplx*>(out)[i] = cadd(yy, cscale(h, s));
}

// ---------------------------------------------------------------------------------------------
// The action ODE's step controller on the device (the host loop it replaces: SciPy's DOP853
// controller as the reference's action integration uses it, src/excitation_action.jl:56-123).
// ---------------------------------------------------------------------------------------------
// rc_ode_step, one workgroup of 256.  finish = 1: judge the attempt that just ran (when active): per
// tracked trajectory its tiles' E5 and E3 sums (slot ranges ranges[tr][r] = (first, count), one wave
// per trajectory, lane-strided with a fixed shuffle tree), |h| e5 / sqrt((e5 + e3 / 100) cnt), the
// worst over trajectories; accept or reject and the next step size.  Then, unless done, set up the
// next attempt: the step clipped at the next stop (a quadrature node, 0 or the window's end) and the
// source coefficients [2|t_s|, s(t_s)] of stages 1..12 (one thread each) with the step at cf[26].
// The controller step shared by rc_ode_step and the persistent launch (c in device or LDS memory; C:
// the error partials are read with device-coherent loads).  No barrier after the cf writes.
template <bool C>
__device__ __forceinline__ void ode_step_body(RcOdeCtl* c, int finish, const double* partial, const int* ranges,
                                              int nranges, int ntrack, double cnt, const double* stops,
                                              const int* stop_node, const double* cheb, double* cf) {
    __shared__ double smax[4];
    __shared__ int sbad[4];
    __shared__ double s_t, s_he, s_tn;
    __shared__ int s_go;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool judge = finish && c->active;   // uniform
    if (judge) {
        const double he = c->he;
        double worst = 0.0;
        int bad = 0;
        for (int tr = wave; tr < ntrack; tr += 4) {
            double e5 = 0.0, e3 = 0.0;
            for (int r = 0; r < nranges; ++r) {
                const int first = ranges[2 * (tr * nranges + r)], count = ranges[2 * (tr * nranges + r) + 1];
                for (int q = lane; q < count; q += 64) {
                    if constexpr (C) {
                        e5 += __hip_atomic_load(partial + 2 * (first + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        e3 += __hip_atomic_load(partial + 2 * (first + q) + 1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
                    } else {
                        e5 += partial[2 * (first + q)];
                        e3 += partial[2 * (first + q) + 1];
                    }
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                e5 += __shfl_xor(e5, off);
                e3 += __shfl_xor(e3, off);
            }
            double et = 0.0;
            if (e5 > 0 || e3 > 0) et = fabs(he) * e5 / sqrt((e5 + 0.01 * e3) * cnt);
            if (!isfinite(et)) bad = 1;
            else worst = fmax(worst, et);
        }
        if (lane == 0) {
            smax[wave] = worst;
            sbad[wave] = bad;
        }
    }
    __syncthreads();
    if (tid == 0) {
        s_go = 0;
        if (judge) {
            const double err = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
            const int bad = sbad[0] | sbad[1] | sbad[2] | sbad[3];
            const double he = c->he;
            c->commit = 0;
            c->sample = -1;
            if (bad) {
                c->status = RC_ST_NONFINITE;
                c->done = 1;
            } else if (err <= 1.0) {
                c->accepted += 1;
                c->commit = 1;
                c->t = c->tn;
                if (c->t == stops[c->si]) {
                    const int nd = stop_node[c->si];
                    if (nd >= 0) {
                        c->sample = nd;
                        c->node = nd + 1;
                    }
                    c->si += 1;
                }
                double factor = err == 0 ? 10.0 : fmin(10.0, 0.9 * pow(err, -1.0 / 8));
                if (c->last_rejected) factor = fmin(1.0, factor);
                const double hn = he * factor;
                c->h = c->clipped ? fmax(hn, c->h) : hn;
                c->last_rejected = 0;
                if (c->si >= c->nstops) {
                    c->done = 1;
                    if (c->node != c->npts) c->status = RC_ST_MISSING_SAMPLES;
                }
            } else {
                c->rejected += 1;
                c->h = he * fmax(0.2, 0.9 * pow(err, -1.0 / 8));
                c->last_rejected = 1;
            }
        }
        c->active = 0;
        if (!c->done) {
            const double t = c->t, h = c->h, stop = stops[c->si];
            double he = fmin(h, stop - t);
            const int clipped = he < h;
            if (stop - t - he <= 1e-14 * fmax(1.0, fabs(stop))) he = stop - t;
            if (!(he > 10 * 2.220446049250313e-16 * fmax(1.0, fabs(t)))) {
                c->status = RC_ST_STEP_UNDERFLOW;
                c->done = 1;
            } else if (++c->attempts > c->maxiters) {
                c->status = RC_ST_MAX_STEPS;
                c->done = 1;
            } else {
                const double tn = (he == stop - t) ? stop : t + he;
                c->he = he;
                c->tn = tn;
                c->clipped = clipped;
                c->active = 1;
                s_t = t;
                s_he = he;
                s_tn = tn;
                s_go = 1;
            }
        }
    }
    __syncthreads();
    if (s_go && tid >= 1 && tid < 13) {
        const double ts = tid < 12 ? s_t + RC_C[tid] * s_he : s_tn;
        cf[2 * tid] = 2 * fabs(ts);
        cf[2 * tid + 1] = rc_source(ts, cheb);
    }
    if (s_go && tid == 0) cf[26] = s_he;
}

__global__ void __launch_bounds__(256) rc_ode_step(RcOdeCtl* c, int finish, const double* partial,
                                                   const int* ranges, int nranges, int ntrack, double cnt,
                                                   const double* stops, const int* stop_node, const double* cheb,
                                                   double* cf) {
    ode_step_body<false>(c, finish, partial, ranges, nranges, ntrack, cnt, stops, stop_node, cheb, cf);
}

// First stage input of the attempt, U1 = y + h a10 K_0, preceded by the commit of the previous
// attempt when it was accepted (y <- y_new, K_0 <- K_12, and at a quadrature node the forward half
// of y_new as sample `sample`, the adjoint half as sample npts - 1 - sample).
__global__ void rc_ode_stage(const RcOdeCtl* c, long long N, long long half, double a10, const double* hp,
                             const double* ynew, double* y, const double* K12, double* K0, double* Sf, double* Sa,
                             double* U1) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int commit = c->commit, active = c->active;
    cplx yi, ki;
    if (commit) {
        yi = reinterpret_cast<const cplx*>(ynew)[i];
        ki = reinterpret_cast<const cplx*>(K12)[i];
        reinterpret_cast<cplx*>(y)[i] = yi;
        reinterpret_cast<cplx*>(K0)[i] = ki;
        const int smp = c->sample;
        if (smp >= 0) {
            if (i < half) reinterpret_cast<cplx*>(Sf)[(long long)smp * half + i] = yi;
            else reinterpret_cast<cplx*>(Sa)[(long long)(c->npts - 1 - smp) * half + (i - half)] = yi;
        }
    } else {
        if (!active) return;
        yi = reinterpret_cast<const cplx*>(y)[i];
        ki = reinterpret_cast<const cplx*>(K0)[i];
    }
    if (active) reinterpret_cast<cplx*>(U1)[i] = cadd(yi, cscale(hp[0], cscale(a10, ki)));
}

// Resolvent step in the transfer's eigenbasis: out[v][k] = Y[src(v)][k] / (1 - z_i (mu_k + i s_b)) for
// vector v = b nz + i (src(v) = b when per_basis, else v), mu conjugated for the adjoint; N entries
// per vector; s null: no shift.  grid (nb nz, ceil(N / 256)).
__global__ void rc_res_scale(long long N, int nz, int per_basis, int adjoint, const double* Y, const double* mu,
                             const double* z, const double* s, double* out) {
    const long long k = (long long)blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int v = blockIdx.x, b = v / nz, i = v % nz;
    const cplx y = reinterpret_cast<const cplx*>(Y)[(long long)(per_basis ? b : v) * N + k];
    cplx l = reinterpret_cast<const cplx*>(mu)[k];
    if (adjoint) l = cconj(l);
    const double zi = z[i], sb = s ? s[b] : 0.0;
    const cplx den = cmake(1.0 - zi * l.re, -zi * (l.im + sb));
    const double m = cabs2(den);
    reinterpret_cast<cplx*>(out)[(long long)v * N + k] =
        cmake((y.re * den.re + y.im * den.im) / m, (y.im * den.re - y.re * den.im) / m);
}

// The pinned Poisson operator in the transfer's eigenbasis: out_k = Y_k / (-l_k + i s + [k == k0]),
// l = mu (adjoint: conj mu), k0 the zero mode; zero on the padding of the Dp^2 space (where the
// eigenvalue is a placeholder 0).  N = Dp^2, grid ceil(N / 256).
__global__ void rc_poisson_scale(long long N, int Dp, int D, int adjoint, double s, int k0, const double* mu,
                                 const double* Y, double* out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    if (k / Dp >= D || k % Dp >= D) {
        reinterpret_cast<cplx*>(out)[k] = cmake(0, 0);
        return;
    }
    cplx l = reinterpret_cast<const cplx*>(mu)[k];
    if (adjoint) l = cconj(l);
    const cplx den = cmake(-l.re + (k == k0 ? 1.0 : 0.0), -l.im + s);
    const cplx y = reinterpret_cast<const cplx*>(Y)[k];
    const double m = cabs2(den);
    reinterpret_cast<cplx*>(out)[k] = cmake((y.re * den.re + y.im * den.im) / m, (y.im * den.re - y.re * den.im) / m);
}

// ---- the pinned Poisson solves and tangent algebra of many tangents at once (blockIdx.y = tangent
// v, strides in complex elements); each kernel repeats its single-matrix counterpart's arithmetic.
// X_v -= t_v Z with t_v the device complex at t + 2 v (rc_sub_scaled with den = 1)
__global__ void rc_sub_scaled_n(int n, const double* t, const double* Z, double* X, long long sx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (i >= n) return;
    const cplx tt = cmake(t[2 * v], t[2 * v + 1]);
    const double dre = 1.0, dim = 0.0, m = dre * dre + dim * dim;
    const cplx f = cmake((tt.re * dre + tt.im * dim) / m, (tt.im * dre - tt.re * dim) / m);
    cplx* x = reinterpret_cast<cplx*>(X) + v * sx;
    x[i] = csub(x[i], cmul(f, reinterpret_cast<const cplx*>(Z)[i]));
}

// X_v += scale t_v I on the D x D block (rc_add_identity)
__global__ void rc_add_identity_n(int Dp, int D, const double* t, double scale, double* X, long long sx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (i >= D) return;
    cplx* x = reinterpret_cast<cplx*>(X) + v * sx;
    x[i * Dp + i].re += scale * t[2 * v];
    x[i * Dp + i].im += scale * t[2 * v + 1];
}

// out + so v = <X_v, X_v>, one workgroup per tangent (Xl = X: the left operand passed apart, so the
// products compile as rc_dots's do)
__global__ void __launch_bounds__(256) rc_norms_n(int n, const double* Xl, const double* X, long long sx, double* out,
                                                  long long so) {
    __shared__ double sre[256], sim[256];
    const int v = blockIdx.x;
    dot_block(reinterpret_cast<const cplx*>(Xl) + v * sx, reinterpret_cast<const cplx*>(X) + v * sx, n, sre, sim,
              out + so * v);
}

// rc_poisson_scale with the shift s_v of tangent v
__global__ void rc_poisson_scale_n(long long N, int Dp, int D, int adjoint, const double* s, int k0, const double* mu,
                                   const double* Y, long long sy, double* out, long long so) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (k >= N) return;
    cplx* o = reinterpret_cast<cplx*>(out) + v * so;
    if (k / Dp >= D || k % Dp >= D) {
        o[k] = cmake(0, 0);
        return;
    }
    cplx l = reinterpret_cast<const cplx*>(mu)[k];
    if (adjoint) l = cconj(l);
    const cplx den = cmake(-l.re + (k == k0 ? 1.0 : 0.0), -l.im + s[v]);
    const cplx y = (reinterpret_cast<const cplx*>(Y) + v * sy)[k];
    const double m = cabs2(den);
    o[k] = cmake((y.re * den.re + y.im * den.im) / m, (y.im * den.re - y.re * den.im) / m);
}

// out_v = a X_v + b_v Y_v with b_v the device complex at b + 2 v (rc_axpby)
__global__ void rc_axpby_n(int n, double ar, double ai, const double* X, long long sx, const double* b, const double* Y,
                           long long sy, double* out, long long so) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (i >= n) return;
    const cplx* x = reinterpret_cast<const cplx*>(X) + v * sx;
    const cplx* y = reinterpret_cast<const cplx*>(Y) + v * sy;
    cplx r = cmul(cmake(ar, ai), x[i]);
    r = cadd(r, cmul(cmake(b[2 * v], b[2 * v + 1]), y[i]));
    (reinterpret_cast<cplx*>(out) + v * so)[i] = r;
}

// X_v += c_v Z_v with c_v the device complex at c + 2 v (rc_axpy_c)
__global__ void rc_axpy_c_n(int n, const double* c, const double* Z, long long sz, double* X, long long sx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (i >= n) return;
    cplx* x = reinterpret_cast<cplx*>(X) + v * sx;
    x[i] = cadd(x[i], cmul(cmake(c[2 * v], c[2 * v + 1]), (reinterpret_cast<const cplx*>(Z) + v * sz)[i]));
}

// out[v] = src[v / nz] (N complex each): block 1 of basis b's right-hand side for each node.
// grid (nb nz, ceil(N / 256)).
__global__ void rc_bcast_rhs(long long N, int nz, const double* src, long long ss, double* out) {
    const long long k = (long long)blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int v = blockIdx.x;
    reinterpret_cast<cplx*>(out)[(long long)v * N + k] = reinterpret_cast<const cplx*>(src)[(long long)(v / nz) * ss + k];
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

// X += c Z with c = -scale sum_j sum_k wz[k] tr[j nodes + k] over nj sets of `nodes` complex traces: the
// sums in increasing j, then k, with each product and sum rounded on its own (the host loop's
// arithmetic), made by every thread alike
__device__ __forceinline__ void trace_axpy_elem(int nj, int nodes, const double* tr, const double* wz, double scale,
                                                const double* Z, double* X, int i) {
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < nj; ++j)
        for (int k = 0; k < nodes; ++k) {
            const double w = wz[k];
            const double* t = tr + 2 * ((long long)j * nodes + k);
            sr = __dadd_rn(sr, __dmul_rn(w, t[0]));
            si = __dadd_rn(si, __dmul_rn(w, t[1]));
        }
    sr = __dmul_rn(sr, scale);
    si = __dmul_rn(si, scale);
    cplx* x = reinterpret_cast<cplx*>(X);
    x[i] = cadd(x[i], cmul(cmake(-sr, -si), reinterpret_cast<const cplx*>(Z)[i]));
}

__global__ void rc_trace_axpy(int n, int nj, int nodes, const double* tr, const double* wz, double scale,
                              const double* Z, double* X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    trace_axpy_elem(nj, nodes, tr, wz, scale, Z, X, i);
}

// rc_trace_axpy for several tangents (blockIdx.y = v): traces tr + 2 v nj nodes, Z + v sz, X + v sx
__global__ void rc_trace_axpy_n(int n, int nj, int nodes, const double* tr, const double* wz, double scale,
                                const double* Z, long long sz, double* X, long long sx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long v = blockIdx.y;
    trace_axpy_elem(nj, nodes, tr + 2 * v * nj * nodes, wz, scale, Z + 2 * v * sz, X + 2 * v * sx, i);
}

// ---------------------------------------------------------------------------------------------
// One right-hand side of the :phi^4: action ODEs (forward and adjoint, every tangent and channel)
// in a single launch.  Workgroup = one 16 x 16 output tile of one derivative matrix; all products
// of that tile accumulate in one MFMA accumulator (k split across the 4 waves), the two-sided
// products A X B through a Dp x 16 column block X B[:, tile] staged in LDS.
//   forward  dx_k = jac (Q x + x Q^+ + R x R^+) + k jt (R x_{k-1} + x_{k-1} R^+)
//            dy_k = jac (Qs y + y Q^+ + R y R^+ + v x + w x R^+) + k jt (R y_{k-1} + y_{k-1} R^+ + w x_{k-1})
//   adjoint  da0_k = jac (Qhs a0 + a0 Q + R^+ (a0 R + a1 w) + a1 v) + (k+1) jt (a0 R + a1 w + R^+ a0)_{k+1}
//            da1_k = jac (Q^+ a1 + a1 Q + R^+ a1 R) + (k+1) jt (a1 R + R^+ a1)_{k+1}
// with Qs = Q - ip, Qhs = Q^+ + ip of each tangent's momentum.  grid ((Dp/16)^2, 4 * 5 na).
// ---------------------------------------------------------------------------------------------
namespace {

// Every fragment of this wave's k-steps (Dp / 16 of them, at most 4) is loaded before the first MFMA,
// so a product pays one load round trip instead of one per k-step; the MFMAs run in the same k order.
// Wave w runs k-step w + 4 i of k-tile i.  kmask (uniform) selects the k-tiles: bit i for k-tile i.
// The stage input as a buffer resource (device-coherent 16-byte loads of the elements another
// workgroup of the same launch wrote; offsets outside it read zero).
struct QCoh {
    __amdgpu_buffer_rsrc_t r;
    const cplx* base;
};

// An operand element; C: a device-coherent load from the stage input q.
template <bool C>
__device__ __forceinline__ cplx qld(const QCoh* q, const cplx* p) {
    if constexpr (C) {
        return ld_coh(q->r, (uint32_t)((p - q->base) * (long long)sizeof(cplx)));
    } else {
        (void)q;
        return *p;
    }
}

template <bool CA = false, bool CB = false>
__device__ __forceinline__ void qprod(cacc& acc, const cplx* A, const cplx* B, int Dp, int bi, int bj, int wave,
                                      int lane, double scale, int kmask = 15, const QCoh* q = nullptr) {
    const int nk = Dp / 16;
    cplx a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nk && ((kmask >> i) & 1)) {
            const int kk = 4 * (wave + 4 * i) + (lane >> 4);
            a[i] = qld<CA>(q, &A[(16 * bi + (lane & 15)) * Dp + kk]);
            b[i] = qld<CB>(q, &B[kk * Dp + 16 * bj + (lane & 15)]);
        }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nk && ((kmask >> i) & 1)) cmfma(acc, cscale(scale, a[i]), b[i]);
}

// A[tile rows, :] T with T a Dp x 16 block in LDS
__device__ __forceinline__ void qprod_t(cacc& acc, const cplx* A, const cplx* T, int Dp, int bi, int wave, int lane,
                                        double scale, int kmask = 15) {
    const int nk = Dp / 16;
    cplx a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nk && ((kmask >> i) & 1)) a[i] = A[(16 * bi + (lane & 15)) * Dp + 4 * (wave + 4 * i) + (lane >> 4)];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nk && ((kmask >> i) & 1)) {
            const int kk = 4 * (wave + 4 * i) + (lane >> 4);
            cmfma(acc, cscale(scale, a[i]), T[kk * 16 + (lane & 15)]);
        }
}

// k-tile mask of A(b, :) B(:, c) for operands of parities pa, pb under parity blocks of zt tiles (zt = 0:
// every k-tile): zero unless the output tile's blocks have parity pa ^ pb, then the k-tiles of block
// blk(b) ^ pa.  The skipped k-steps multiply exact zeros, so the sums keep their bits.
__device__ __forceinline__ int zmask(int zt, int nk, int b, int c, int pa, int pb) {
    if (zt == 0) return 15;
    const int bb = b / zt, bc = c / zt;
    if ((bb ^ bc) != (pa ^ pb)) return 0;
    int m = 0;
    for (int t = 0; t < nk; ++t)
        if (t / zt == (bb ^ pa)) m |= 1 << t;
    return m;
}

__device__ __forceinline__ cplx qreduce(const cacc& acc, cplx* red, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 256 + (4 * q + (lane >> 4)) * 16 + (lane & 15)] = cacc_get(acc, q);
    __syncthreads();
    const int t = threadIdx.x;
    const cplx s = cadd(cadd(cadd(red[t], red[256 + t]), red[512 + t]), red[768 + t]);
    __syncthreads();
    return s;
}

// One wave's whole product A(bi, :) B(:, bj) over every k-step (no cross-wave reduction), one k-tile of
// fragments loaded ahead of its four MFMAs; kmask selects the k-tiles (bit t: k-steps 4t .. 4t + 3).
template <bool CA = false>
__device__ __forceinline__ void qprod_rows(cacc& acc, const cplx* A, const cplx* B, int Dp, int bi, int bj, int lane,
                                           int kmask, const QCoh* q = nullptr) {
    const int nk = Dp / 16;
    for (int t = 0; t < nk; ++t) {
        if (!((kmask >> t) & 1)) continue;
        cplx a[4], b[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = 16 * t + 4 * s + (lane >> 4);
            a[s] = qld<CA>(q, &A[(16 * bi + (lane & 15)) * Dp + kk]);
            b[s] = B[kk * Dp + 16 * bj + (lane & 15)];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) cmfma(acc, a[s], b[s]);
    }
}

// one term X B[:, bj] of an LDS column block, with the operands' Z2 parities (X null: no term)
struct QTerm {
    const cplx* X;
    const cplx* B;
    int pa, pb;
};

// Column blocks (Dp x 16, in LDS) T0 = a0 + a1 and, when T1 is given, T1 = b0 + b1.  The row tiles of
// both blocks are spread over the waves, each wave running every k-step of its row tiles and writing
// them straight to LDS; one barrier at the end.  zt: Z2 parity blocks (0: every k-tile).
template <bool CX = false>
__device__ __forceinline__ void qcolumns(cplx* T0, QTerm a0, QTerm a1, cplx* T1, QTerm b0, QTerm b1, int Dp, int bj,
                                         int wave, int lane, int zt, const QCoh* qc = nullptr) {
    const int nbt = Dp / 16, ntask = (T1 ? 2 : 1) * nbt;
    for (int i = wave; i < ntask; i += 4) {
        const int col = i / nbt, rt = i % nbt;
        const QTerm p = col ? b0 : a0, q = col ? b1 : a1;
        cplx* T = col ? T1 : T0;
        cacc acc;
        cacc_zero(acc);
        qprod_rows<CX>(acc, p.X, p.B, Dp, rt, bj, lane, zmask(zt, nbt, rt, bj, p.pa, p.pb), qc);
        if (q.X) qprod_rows<CX>(acc, q.X, q.B, Dp, rt, bj, lane, zmask(zt, nbt, rt, bj, q.pa, q.pb), qc);
#pragma unroll
        for (int e = 0; e < 4; ++e) T[(16 * rt + 4 * e + (lane >> 4)) * 16 + (lane & 15)] = cacc_get(acc, e);
    }
    __syncthreads();
}

}  // namespace

// One right-hand-side tile of the :phi^4: action ODE (the products of rc_quartic_rhs for output tile
// (bi, bj) of matrix mi = blockIdx.y, reduced over the waves): thread t receives element (t/16, t%16).
// U: the stage input; C: its elements are read with device-coherent loads (another workgroup of the
// same launch wrote them).
template <bool C>
__device__ __forceinline__ cplx quartic_tile(const RcQuarticArgs& g, const cplx* U, double jac, double jt, int bi,
                                             int bj, int mi, cplx* red, cplx* T1, cplx* T2) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int Dp = g.Dp;
    const int nb = g.na * RC_QCH;
    // block, tangent and channel of matrix mi (rc_qmat's inverse; the shared blocks 0 and 3 take a = 0)
    int blk, a, k;
    if (mi < RC_QCH) { blk = 0; a = 0; k = mi; }
    else if (mi < RC_QCH + nb) { blk = 1; a = (mi - RC_QCH) / RC_QCH; k = (mi - RC_QCH) % RC_QCH; }
    else if (mi < RC_QCH + 2 * nb) { blk = 2; a = (mi - RC_QCH - nb) / RC_QCH; k = (mi - RC_QCH - nb) % RC_QCH; }
    else { blk = 3; a = 0; k = mi - RC_QCH - 2 * nb; }
    const long long m = (long long)Dp * Dp;
    auto mat = [&](int b, int ch) { return U + (long long)rc_qmat(b, a, ch, nb) * m; };
    const cplx* Q = reinterpret_cast<const cplx*>(g.Q);
    const cplx* Qh = reinterpret_cast<const cplx*>(g.Qh);
    const cplx* Qs = reinterpret_cast<const cplx*>(g.Qs) + (long long)a * g.qstride;     // tangent's momentum
    const cplx* Qhs = reinterpret_cast<const cplx*>(g.Qhs) + (long long)a * g.qstride;
    const cplx* R = reinterpret_cast<const cplx*>(g.R);
    const cplx* Rh = reinterpret_cast<const cplx*>(g.Rh);
    const cplx* v = reinterpret_cast<const cplx*>(g.Vrep) + (long long)a * 5 * m;
    const cplx* w = reinterpret_cast<const cplx*>(g.Wrep) + (long long)a * 5 * m;
    // Z2 parities (g.zt > 0): Q, Q^+ and the shifted ones even, R, R^+ odd, w = W of parity pw,
    // v = -R^+ W of parity pw ^ 1; channels x_k and a1_k of parity k & 1 (x_0 = rho, a1 of the last
    // channel = I), y_k and a0_k of parity (k + 1 + pw) & 1 (y_0 = R W, a0 driven by a1 w and a1 v)
    const int zt = g.zt, nk = Dp / 16, pw = g.wpar;
    const int px = k & 1, py = (k + 1 + pw) & 1;
    auto zm = [&](int pa, int pb) { return zmask(zt, nk, bi, bj, pa, pb); };
    const QCoh qc{rc_rsrc(U, (uint32_t)((long long)rc_qmats(nb) * m * (long long)sizeof(cplx))), U};
    cacc acc;
    cacc_zero(acc);
    if (blk == 0) {
        const cplx* x = mat(0, k);
        if (jac != 0.0) {
            qprod<false, C>(acc, Q, x, Dp, bi, bj, wave, lane, jac, zm(0, px), &qc);
            qprod<C, false>(acc, x, Qh, Dp, bi, bj, wave, lane, jac, zm(px, 0), &qc);
            qcolumns<C>(T1, QTerm{x, Rh, px, 1}, QTerm{}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, zt, &qc);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac, zm(1, px ^ 1));
        }
        if (k >= 1) {
            const cplx* xp = mat(0, k - 1);
            qprod<false, C>(acc, R, xp, Dp, bi, bj, wave, lane, k * jt, zm(1, px ^ 1), &qc);
            qprod<C, false>(acc, xp, Rh, Dp, bi, bj, wave, lane, k * jt, zm(px ^ 1, 1), &qc);
        }
    } else if (blk == 1) {
        const cplx* y = mat(1, k);
        const cplx* x = mat(0, k);
        if (jac != 0.0) {
            qprod<false, C>(acc, Qs, y, Dp, bi, bj, wave, lane, jac, zm(0, py), &qc);
            qprod<C, false>(acc, y, Qh, Dp, bi, bj, wave, lane, jac, zm(py, 0), &qc);
            qprod<false, C>(acc, v, x, Dp, bi, bj, wave, lane, jac, zm(pw ^ 1, px), &qc);
            qcolumns<C>(T1, QTerm{y, Rh, py, 1}, QTerm{}, T2, QTerm{x, Rh, px, 1}, QTerm{}, Dp, bj, wave, lane, zt, &qc);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac, zm(1, py ^ 1));
            qprod_t(acc, w, T2, Dp, bi, wave, lane, jac, zm(pw, px ^ 1));
        }
        if (k >= 1) {
            const cplx* yp = mat(1, k - 1);
            const cplx* xp = mat(0, k - 1);
            qprod<false, C>(acc, R, yp, Dp, bi, bj, wave, lane, k * jt, zm(1, py ^ 1), &qc);
            qprod<C, false>(acc, yp, Rh, Dp, bi, bj, wave, lane, k * jt, zm(py ^ 1, 1), &qc);
            qprod<false, C>(acc, w, xp, Dp, bi, bj, wave, lane, k * jt, zm(pw, px ^ 1), &qc);
        }
    } else if (blk == 2) {
        const cplx* a0 = mat(2, k);
        const cplx* a1 = mat(3, k);
        if (jac != 0.0) {
            qprod<false, C>(acc, Qhs, a0, Dp, bi, bj, wave, lane, jac, zm(0, py), &qc);
            qprod<C, false>(acc, a0, Q, Dp, bi, bj, wave, lane, jac, zm(py, 0), &qc);
            qprod<C, false>(acc, a1, v, Dp, bi, bj, wave, lane, jac, zm(px, pw ^ 1), &qc);
            qcolumns<C>(T1, QTerm{a0, R, py, 1}, QTerm{a1, w, px, pw}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, zt, &qc);
            qprod_t(acc, Rh, T1, Dp, bi, wave, lane, jac, zm(1, py ^ 1));
        }
        if (k <= 3) {
            const double c = (k + 1) * jt;
            const cplx* a0n = mat(2, k + 1);
            const cplx* a1n = mat(3, k + 1);
            qprod<C, false>(acc, a0n, R, Dp, bi, bj, wave, lane, c, zm(py ^ 1, 1), &qc);
            qprod<C, false>(acc, a1n, w, Dp, bi, bj, wave, lane, c, zm(px ^ 1, pw), &qc);
            qprod<false, C>(acc, Rh, a0n, Dp, bi, bj, wave, lane, c, zm(1, py ^ 1), &qc);
        }
    } else {
        const cplx* a1 = mat(3, k);
        if (jac != 0.0) {
            qprod<false, C>(acc, Qh, a1, Dp, bi, bj, wave, lane, jac, zm(0, px), &qc);
            qprod<C, false>(acc, a1, Q, Dp, bi, bj, wave, lane, jac, zm(px, 0), &qc);
            qcolumns<C>(T1, QTerm{a1, R, px, 1}, QTerm{}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, zt, &qc);
            qprod_t(acc, Rh, T1, Dp, bi, wave, lane, jac, zm(1, px ^ 1));
        }
        if (k <= 3) {
            const double c = (k + 1) * jt;
            const cplx* a1n = mat(3, k + 1);
            qprod<C, false>(acc, a1n, R, Dp, bi, bj, wave, lane, c, zm(px ^ 1, 1), &qc);
            qprod<false, C>(acc, Rh, a1n, Dp, bi, bj, wave, lane, c, zm(1, px ^ 1), &qc);
        }
    }
    (void)tid;
    return qreduce(acc, red, wave, lane);
}

__global__ void __launch_bounds__(256) rc_quartic_rhs(RcQuarticArgs g) {
    __shared__ cplx red[1024];
    __shared__ cplx T1[64 * 16];
    __shared__ cplx T2[64 * 16];
    if (g.active && !*g.active) return;
    const int tid = threadIdx.x;
    const int Dp = g.Dp, nbt = Dp / 16, bi = blockIdx.x / nbt, bj = blockIdx.x % nbt;
    const int mi = blockIdx.y;
    const long long m = (long long)Dp * Dp;
    const cplx* U = reinterpret_cast<const cplx*>(g.U);
    const double jac = g.coef ? g.coef[0] : g.jac;
    const double jt = g.coef ? g.coef[1] : g.jt;
    const cplx val = quartic_tile<false>(g, U, jac, jt, bi, bj, mi, red, T1, T2);
    const long long e = (long long)mi * m + (16 * bi + (tid >> 4)) * Dp + 16 * bj + (tid & 15);
    reinterpret_cast<cplx*>(g.F)[e] = val;
    if (g.epi == 0) return;
    const cplx* Kc = reinterpret_cast<const cplx*>(g.K);
    const cplx y0 = reinterpret_cast<const cplx*>(g.y)[e];
    if (g.epi == 1) {
        const double h = g.hp ? g.hp[0] : g.h;
        cplx sum = cscale(g.comb.a[g.stage], val);
        for (int j = 0; j < g.stage; ++j)
            if (g.comb.a[j] != 0.0) sum = cadd(sum, cscale(g.comb.a[j], Kc[(long long)j * g.sk + e]));
        reinterpret_cast<cplx*>(g.Unext)[e] = cadd(y0, cscale(h, sum));
        return;
    }
    cplx r5 = cscale(g.comb.a[g.stage], val), r3 = cscale(g.comb3.a[g.stage], val);
    for (int j = 0; j < g.stage; ++j) {
        const cplx kj = Kc[(long long)j * g.sk + e];
        if (g.comb.a[j] != 0.0) r5 = cadd(r5, cscale(g.comb.a[j], kj));
        if (g.comb3.a[j] != 0.0) r3 = cadd(r3, cscale(g.comb3.a[j], kj));
    }
    const double sc = g.atol + g.rtol * fmax(cabsd(y0), cabsd(U[e]));
    double* s5 = reinterpret_cast<double*>(red);
    double* s3 = s5 + 256;
    s5[tid] = cabs2(r5) / (sc * sc);
    s3[tid] = cabs2(r3) / (sc * sc);
    __syncthreads();
    for (int hh = 128; hh > 0; hh >>= 1) {
        if (tid < hh) { s5[tid] += s5[tid + hh]; s3[tid] += s3[tid + hh]; }
        __syncthreads();
    }
    if (tid == 0) {
        const long long slot = (long long)mi * nbt * nbt + blockIdx.x;
        g.partial[2 * slot] = s5[0];
        g.partial[2 * slot + 1] = s3[0];
    }
}

// ---------------------------------------------------------------------------------------------
// The whole :phi^4: action ODE integration in one persistent launch (cooperative: every workgroup
// resident).  Workgroup (bi, bj, mi) owns the same 16 x 16 tile as in rc_quartic_rhs and keeps its
// elements of y, of the stage derivatives K_0..K_12 (LDS) and of the step end in registers.  Each
// stage input is exchanged through `ubuf` (two buffers, device-coherent stores and loads) with one
// all-workgroup event per stage: every workgroup publishes its counter and waits for all others.  The
// step controller runs redundantly in every workgroup on the same error partials (exchanged the same
// way), so every workgroup takes the same decisions; the arithmetic is the graph path's, so the
// results are the same bits.  Waits are bounded in wall time: a workgroup that outlasts the budget
// sets the abort word and every workgroup leaves with RC_ST_WAIT_LIMIT.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rc_quartic_ode(const RcOdeRun* __restrict__ pp) {
    __shared__ cplx red[1024];
    __shared__ cplx T1[64 * 16];
    __shared__ cplx T2[64 * 16];
    __shared__ cplx sK[13][256];
    __shared__ RcOdeCtl sc;
    __shared__ double s_cf[32];
    __shared__ int s_abort;
    const RcOdeRun& p = *pp;
    const RcQuarticArgs& g = p.g;
    const int tid = threadIdx.x, lane = tid & 63;
    const int Dp = g.Dp, nbt = Dp / 16, bi = blockIdx.x / nbt, bj = blockIdx.x % nbt, mi = blockIdx.y;
    const int wg = mi * nbt * nbt + (int)blockIdx.x;   // also the tile's error slot
    const long long m = (long long)Dp * Dp;
    const long long e = (long long)mi * m + (16 * bi + (tid >> 4)) * Dp + 16 * bj + (tid & 15);
    cplx* const u0 = reinterpret_cast<cplx*>(p.u0);
    cplx* const u1 = reinterpret_cast<cplx*>(p.u1);
    uint32_t ev = 0;
    auto put = [&](int buf, cplx v) {
        double* d = reinterpret_cast<double*>((buf ? u1 : u0) + e);
        __hip_atomic_store(d, v.re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(d + 1, v.im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // publish this workgroup's next event once its stores are complete, then wait until every
    // workgroup has published it (wave 0 polls every counter in each round: lane l those of
    // workgroups l, l + 64, l + 128, l + 192; at most 256 workgroups)
    auto sync_all = [&]() -> bool {
        ++ev;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) st_flag(p.flags + wg, ev);
        if (tid < 64) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            int bad = 0;
            for (uint32_t polls = 1;; ++polls) {
                uint32_t fv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) fv[q] = lane + 64 * q < p.nwg ? ld_flag(p.flags + lane + 64 * q) : ev;
                const uint32_t ab = ld_flag(p.abort_flag);
                if (__all(fv[0] >= ev && fv[1] >= ev && fv[2] >= ev && fv[3] >= ev)) break;
                if (ab) { bad = 1; break; }
                if ((polls & 63) == 0 && (long long)(__builtin_amdgcn_s_memrealtime() - t0) > p.wait_ticks) {
                    bad = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (lane == 0) {
                if (bad) st_flag(p.abort_flag, 1u);
                s_abort = bad;
            }
        }
        __syncthreads();
        return s_abort == 0;
    };
    // workgroup 0's time (constant-rate clock) in the products, the exchanges and the controller
    uint64_t t_rhs = 0, t_sync = 0, t_ctl = 0, tk = __builtin_amdgcn_s_memrealtime();
    const uint64_t t_start = tk;
    auto lap = [&](uint64_t& acc) {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        acc += now - tk;
        tk = now;
    };
    cplx yv = reinterpret_cast<const cplx*>(p.ybuf)[e];
    if (tid == 0) sc = p.c0;
    bool ok = true;
    // the first right-hand side (first same as last) on y, exchanged through buffer 1 (the first
    // attempt writes buffer 0 before its event, buffer 1 only after the next)
    put(1, yv);
    if (!(ok = sync_all())) goto done;
    sK[0][tid] = quartic_tile<true>(g, u1, p.jac0, p.jt0, bi, bj, mi, red, T1, T2);
    ode_step_body<true>(&sc, 0, p.slots, p.ranges, p.nranges, p.ntrack, p.cnt, p.stops, p.stop_node, p.cheb, s_cf);
    __syncthreads();
    while (sc.active) {
        const double he = s_cf[26];
        put(0, cadd(yv, cscale(he, cscale(RC_A[1][0], sK[0][tid]))));
        lap(t_ctl);
        if (!(ok = sync_all())) break;
        lap(t_sync);
        cplx ynew = cmake(0, 0);
        for (int st = 1; st < 13; ++st) {
            const cplx F = quartic_tile<true>(g, ((st - 1) & 1) ? u1 : u0, s_cf[2 * st], s_cf[2 * st + 1], bi, bj,
                                              mi, red, T1, T2);
            lap(t_rhs);
            sK[st][tid] = F;
            if (st < 12) {   // the next stage input (stage 12's is the step end, B weights)
                cplx sum = cscale(st < 11 ? RC_A[st + 1][st] : RC_B[st], F);
                for (int j = 0; j < st; ++j) {
                    const double aj = st < 11 ? RC_A[st + 1][j] : RC_B[j];
                    if (aj != 0.0) sum = cadd(sum, cscale(aj, sK[j][tid]));
                }
                const cplx u = cadd(yv, cscale(he, sum));
                if (st == 11) ynew = u;
                put(st & 1, u);
                if (!(ok = sync_all())) break;
                lap(t_sync);
            } else {         // error sums of this tile
                cplx r5 = cscale(RC_E5[12], F), r3 = cscale(RC_E3[12], F);
                for (int j = 0; j < 12; ++j) {
                    const cplx kj = sK[j][tid];
                    if (RC_E5[j] != 0.0) r5 = cadd(r5, cscale(RC_E5[j], kj));
                    if (RC_E3[j] != 0.0) r3 = cadd(r3, cscale(RC_E3[j], kj));
                }
                const double scl = p.atol + p.rtol * fmax(cabsd(yv), cabsd(ynew));
                double* s5 = reinterpret_cast<double*>(red);
                double* s3 = s5 + 256;
                s5[tid] = cabs2(r5) / (scl * scl);
                s3[tid] = cabs2(r3) / (scl * scl);
                __syncthreads();
                for (int hh = 128; hh > 0; hh >>= 1) {
                    if (tid < hh) { s5[tid] += s5[tid + hh]; s3[tid] += s3[tid + hh]; }
                    __syncthreads();
                }
                if (tid == 0) {
                    __hip_atomic_store(p.slots + 2 * wg, s5[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(p.slots + 2 * wg + 1, s3[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (!(ok = sync_all())) break;
                lap(t_sync);
            }
        }
        if (!ok) break;
        ode_step_body<true>(&sc, 1, p.slots, p.ranges, p.nranges, p.ntrack, p.cnt, p.stops, p.stop_node, p.cheb,
                            s_cf);
        __syncthreads();
        if (sc.commit) {
            yv = ynew;
            sK[0][tid] = sK[12][tid];
            const int smp = sc.sample;
            if (smp >= 0) {
                if (e < p.half) reinterpret_cast<cplx*>(p.Sf)[(long long)smp * p.half + e] = ynew;
                else reinterpret_cast<cplx*>(p.Sa)[(long long)(p.c0.npts - 1 - smp) * p.half + (e - p.half)] = ynew;
            }
        }
    }
done:
    reinterpret_cast<cplx*>(p.ybuf)[e] = yv;
    if (wg == 0 && tid == 0) {
        p.out[0] = ok ? (double)sc.status : (double)RC_ST_WAIT_LIMIT;
        p.out[1] = (double)sc.attempts;
        p.out[2] = (double)sc.accepted;
        p.out[3] = (double)sc.rejected;
        p.out[4] = (double)t_rhs;
        p.out[5] = (double)t_sync;
        p.out[6] = (double)t_ctl;
        p.out[7] = (double)(__builtin_amdgcn_s_memrealtime() - t_start);
    }
}

// workgroups of rc_quartic_ode that can be resident at once on the current device
hipError_t rc_quartic_ode_capacity(int* blocks) {
    int per_cu = 0, dev = 0, cus = 0, coop = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)rc_quartic_ode, 256, 0)) != hipSuccess)
        return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    *blocks = coop ? per_cu * cus : 0;
    return hipSuccess;
}

hipError_t rc_launch_quartic_ode(dim3 grid, const RcOdeRun* dparams, hipStream_t stream) {
    void* args[] = {(void*)&dparams};
    return hipLaunchCooperativeKernel((const void*)rc_quartic_ode, grid, dim3(256), args, 0, stream);
}

// ---------------------------------------------------------------------------------------------
// L(X) = Q X + X Q^+ + R X R^+ (adjoint: L*(X) = Q^+ X + X Q + R^+ X R, src/transfer.jl:2-19) of
// one Dp x Dp matrix in one launch: one workgroup per 16 x 16 output tile, the one-sided products
// k-split over the waves into one accumulator, the two-sided term through the LDS column block
// T = X R^+[:, bj] (adjoint X R[:, bj]).  Dp <= 64; out must not alias X.  grid (Dp/16)^2.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void lindblad_tile(int Dp, int adjoint, const double* Q, const double* Qh, const double* R,
                                              const double* Rh, const double* X, double* out, cplx* red, cplx* T) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbt = Dp / 16, bi = blockIdx.x / nbt, bj = blockIdx.x % nbt;
    const cplx* x = reinterpret_cast<const cplx*>(X);
    const cplx* a1 = reinterpret_cast<const cplx*>(adjoint ? Qh : Q);    // a1 X
    const cplx* b2 = reinterpret_cast<const cplx*>(adjoint ? Q : Qh);    // X b2
    const cplx* c = reinterpret_cast<const cplx*>(adjoint ? R : Rh);     // T = X c[:, bj]
    const cplx* l = reinterpret_cast<const cplx*>(adjoint ? Rh : R);     // l T
    qcolumns(T, QTerm{x, c, 0, 0}, QTerm{}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, 0);
    cacc acc;
    cacc_zero(acc);
    qprod(acc, a1, x, Dp, bi, bj, wave, lane, 1.0);
    qprod(acc, x, b2, Dp, bi, bj, wave, lane, 1.0);
    qprod_t(acc, l, T, Dp, bi, wave, lane, 1.0);
    const cplx v = qreduce(acc, red, wave, lane);
    reinterpret_cast<cplx*>(out)[(long long)(16 * bi + (tid >> 4)) * Dp + 16 * bj + (tid & 15)] = v;
}

__global__ void __launch_bounds__(256) rc_lindblad(int Dp, int adjoint, const double* Q, const double* Qh,
                                                   const double* R, const double* Rh, const double* X,
                                                   double* out) {
    __shared__ cplx red[1024];
    __shared__ cplx T[64 * 16];
    lindblad_tile(Dp, adjoint, Q, Qh, R, Rh, X, out, red, T);
}

// rc_lindblad on count operands X + v sx into out + v so (blockIdx.y = v), one direction
__global__ void __launch_bounds__(256) rc_lindblad_n(int Dp, int adjoint, const double* Q, const double* Qh,
                                                     const double* R, const double* Rh, const double* X, long long sx,
                                                     double* out, long long so) {
    __shared__ cplx red[1024];
    __shared__ cplx T[64 * 16];
    const long long v = blockIdx.y;
    lindblad_tile(Dp, adjoint, Q, Qh, R, Rh, X + 2 * v * sx, out + 2 * v * so, red, T);
}

// rc_lindblad on up to four operands at once (blockIdx.y = operand), each with its own direction
__global__ void __launch_bounds__(256) rc_lindblad_multi(int Dp, const double* Q, const double* Qh, const double* R,
                                                         const double* Rh, RcLindbladMulti M) {
    __shared__ cplx red[1024];
    __shared__ cplx T[64 * 16];
    const int b = blockIdx.y;
    lindblad_tile(Dp, M.adj[b], Q, Qh, R, Rh, M.X[b], M.out[b], red, T);
}

// ---------------------------------------------------------------------------------------------
// out = A X B in one launch (one workgroup per 16 x 16 output tile, X B[:, bj] as an LDS column
// block), optionally followed by the no-jump inverse's elementwise step in the eigenbasis of Q:
// out_ij <- -out_ij / den_ij, den = lam_i + conj(lam_j) - s (adjoint: conj(lam_i) + lam_j - s),
// s = s_re + i s_im, zero outside the D x D block.  Dp <= 64; out must not alias X.  grid (Dp/16)^2.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rc_two_sided(int Dp, int D, const double* A, const double* X, const double* B,
                                                    double* out, int sylvester, int adjoint, const double* lam,
                                                    double s_re, double s_im) {
    __shared__ cplx red[1024];
    __shared__ cplx T[64 * 16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbt = Dp / 16, bi = blockIdx.x / nbt, bj = blockIdx.x % nbt;
    qcolumns(T, QTerm{reinterpret_cast<const cplx*>(X), reinterpret_cast<const cplx*>(B), 0, 0}, QTerm{}, nullptr,
             QTerm{}, QTerm{}, Dp, bj, wave, lane, 0);
    cacc acc;
    cacc_zero(acc);
    qprod_t(acc, reinterpret_cast<const cplx*>(A), T, Dp, bi, wave, lane, 1.0);
    cplx v = qreduce(acc, red, wave, lane);
    const int r = 16 * bi + (tid >> 4), c = 16 * bj + (tid & 15);
    if (sylvester) {
        if (r >= D || c >= D) {
            v = cmake(0, 0);
        } else {
            const cplx* l = reinterpret_cast<const cplx*>(lam);
            cplx den = adjoint ? cadd(cconj(l[r]), l[c]) : cadd(l[r], cconj(l[c]));
            den.re -= s_re;
            den.im -= s_im;
            const double m = cabs2(den);
            v = cmake(-(v.re * den.re + v.im * den.im) / m, -(v.im * den.re - v.re * den.im) / m);
        }
    }
    reinterpret_cast<cplx*>(out)[(long long)r * Dp + c] = v;
}

// ---------------------------------------------------------------------------------------------
// The connected four-block transfer of a tangent pair (u, w) (blocks 00, ket, bra, ket+bra) applied
// to nch channels in one launch: output matrix (pair p, channel c, block k) =
//   jac T(x_c)_k + c jt S(x_{c-1})_k
// with T the four-block transfer (L on every block, -ip on the ket block, +ip on the bra block, the
// ket insertion v x + w x R^+ from blocks 0 -> 1 and 2 -> 3, the bra insertion x vbra^+ + R x u^+ from
// blocks 0 -> 2 and 1 -> 3, and w x0 u^+ - n x0 into block 3) and S the :phi^4: edge sources (left:
// R x_k, w x_0 -> 1, w x_2 -> 3; right: x_k R^+, x_0 u^+ -> 2, x_1 u^+ -> 3).  With nch = 1, jac = 1 and
// jt = 0 this is the transfer alone (the Krylov operator of the kinetic resolvent).  grid
// ((Dp/16)^2, npair * nch * 4).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rc_pair_rhs(RcPairArgs g) {
    __shared__ cplx red[1024];
    __shared__ cplx T1[64 * 16];
    __shared__ cplx T2[64 * 16];
    if (g.active && !*g.active) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int Dp = g.Dp, nbt = Dp / 16, bi = blockIdx.x / nbt, bj = blockIdx.x % nbt;
    const int per = 4 * g.nch, mi = blockIdx.y, p = mi / per, r = mi % per, c = r / 4, k = r % 4;
    const long long m = (long long)Dp * Dp;
    const cplx* U = reinterpret_cast<const cplx*>(g.U) + (long long)p * g.ustride;
    auto x = [&](int ch, int blk) { return U + (long long)(4 * ch + blk) * m; };
    const long long dofs = (long long)p * g.dstride;
    const cplx* Q = reinterpret_cast<const cplx*>(g.Q);
    const cplx* Qh = reinterpret_cast<const cplx*>(g.Qh);
    const cplx* R = reinterpret_cast<const cplx*>(g.R);
    const cplx* Rh = reinterpret_cast<const cplx*>(g.Rh);
    const cplx* Qs = reinterpret_cast<const cplx*>(g.Qs) + dofs;
    const cplx* Qb = reinterpret_cast<const cplx*>(g.Qb) + dofs;
    const cplx* v = reinterpret_cast<const cplx*>(g.V) + dofs;
    const cplx* w = reinterpret_cast<const cplx*>(g.W) + dofs;
    const cplx* vbh = reinterpret_cast<const cplx*>(g.Vbh) + dofs;
    const cplx* uh = reinterpret_cast<const cplx*>(g.Uh) + dofs;
    const double jac = g.coef ? g.coef[0] : g.jac;
    const double jt = g.coef ? g.coef[1] : g.jt;
    const double cj = c * jt;
    cacc acc;
    cacc_zero(acc);
    if (jac != 0.0) {
        const cplx* xk = x(c, k);
        qprod(acc, k == 1 ? Qs : k == 2 ? Qb : Q, xk, Dp, bi, bj, wave, lane, jac);
        qprod(acc, xk, Qh, Dp, bi, bj, wave, lane, jac);
        if (k == 0) {
            qcolumns(T1, QTerm{xk, Rh, 0, 0}, QTerm{}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, 0);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac);
        } else if (k == 1) {
            qprod(acc, v, x(c, 0), Dp, bi, bj, wave, lane, jac);
            qcolumns(T1, QTerm{xk, Rh, 0, 0}, QTerm{}, T2, QTerm{x(c, 0), Rh, 0, 0}, QTerm{}, Dp, bj, wave, lane, 0);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac);
            qprod_t(acc, w, T2, Dp, bi, wave, lane, jac);
        } else if (k == 2) {
            qprod(acc, x(c, 0), vbh, Dp, bi, bj, wave, lane, jac);
            qcolumns(T1, QTerm{xk, Rh, 0, 0}, QTerm{x(c, 0), uh, 0, 0}, nullptr, QTerm{}, QTerm{}, Dp, bj, wave, lane, 0);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac);
        } else {
            qprod(acc, v, x(c, 2), Dp, bi, bj, wave, lane, jac);
            qprod(acc, x(c, 1), vbh, Dp, bi, bj, wave, lane, jac);
            qcolumns(T1, QTerm{xk, Rh, 0, 0}, QTerm{x(c, 1), uh, 0, 0}, T2, QTerm{x(c, 2), Rh, 0, 0},
                     QTerm{x(c, 0), uh, 0, 0}, Dp, bj, wave, lane, 0);
            qprod_t(acc, R, T1, Dp, bi, wave, lane, jac);
            qprod_t(acc, w, T2, Dp, bi, wave, lane, jac);
        }
    }
    if (c >= 1 && cj != 0.0) {
        const cplx* xp = x(c - 1, k);
        qprod(acc, R, xp, Dp, bi, bj, wave, lane, cj);
        qprod(acc, xp, Rh, Dp, bi, bj, wave, lane, cj);
        if (k == 1) qprod(acc, w, x(c - 1, 0), Dp, bi, bj, wave, lane, cj);
        if (k == 2) qprod(acc, x(c - 1, 0), uh, Dp, bi, bj, wave, lane, cj);
        if (k == 3) {
            qprod(acc, w, x(c - 1, 2), Dp, bi, bj, wave, lane, cj);
            qprod(acc, x(c - 1, 1), uh, Dp, bi, bj, wave, lane, cj);
        }
    }
    cplx val = qreduce(acc, red, wave, lane);
    const long long local = (long long)(16 * bi + (tid >> 4)) * Dp + 16 * bj + (tid & 15);
    if (k == 3 && jac != 0.0) {   // - n x0
        const cplx nn = cmake(g.nov[2 * (long long)p * g.nstride], g.nov[2 * (long long)p * g.nstride + 1]);
        val = csub(val, cscale(jac, cmul(nn, x(c, 0)[local])));
    }
    const long long e = (long long)p * g.fstride + (long long)r * m + local;
    reinterpret_cast<cplx*>(g.F)[e] = val;
    if (g.epi == 0) return;
    // the ODE layout is contiguous (ustride = fstride = 4 nch Dp^2): e indexes y, K and Unext too
    const cplx* Kc = reinterpret_cast<const cplx*>(g.K);
    const cplx y0 = reinterpret_cast<const cplx*>(g.y)[e];
    if (g.epi == 1) {
        const double h = g.hp ? g.hp[0] : g.h;
        cplx sum = cscale(g.comb.a[g.stage], val);
        for (int j = 0; j < g.stage; ++j)
            if (g.comb.a[j] != 0.0) sum = cadd(sum, cscale(g.comb.a[j], Kc[(long long)j * g.sk + e]));
        reinterpret_cast<cplx*>(g.Unext)[e] = cadd(y0, cscale(h, sum));
        return;
    }
    cplx r5 = cscale(g.comb.a[g.stage], val), r3 = cscale(g.comb3.a[g.stage], val);
    for (int j = 0; j < g.stage; ++j) {
        const cplx kj = Kc[(long long)j * g.sk + e];
        if (g.comb.a[j] != 0.0) r5 = cadd(r5, cscale(g.comb.a[j], kj));
        if (g.comb3.a[j] != 0.0) r3 = cadd(r3, cscale(g.comb3.a[j], kj));
    }
    const double sc = g.atol + g.rtol * fmax(cabsd(y0), cabsd(reinterpret_cast<const cplx*>(g.U)[e]));
    double* s5 = reinterpret_cast<double*>(red);
    double* s3 = s5 + 256;
    s5[tid] = cabs2(r5) / (sc * sc);
    s3[tid] = cabs2(r3) / (sc * sc);
    __syncthreads();
    for (int hh = 128; hh > 0; hh >>= 1) {
        if (tid < hh) { s5[tid] += s5[tid + hh]; s3[tid] += s3[tid + hh]; }
        __syncthreads();
    }
    if (tid == 0) {
        const long long slot = (long long)mi * nbt * nbt + blockIdx.x;
        g.partial[2 * slot] = s5[0];
        g.partial[2 * slot + 1] = s3[0];
    }
}
