// Persistent dataflow DOP853 for the six reduced transfer-ODE systems of fg_phi4
// (kinetic, kinetic commutator, phi^4) x (forward, adjoint), FP64 on MFMA.
//
// Equations (reduced channels of src/observables.jl:37-63; Lindblad maps src/transfer.jl:2-19):
//   forward  k(X) = jac(t) [Q X + X Q^+ + (R X) R^+] + s(t) [insertions + fixed]
//   adjoint  k(X) = jac(t) [Q^+ X + X Q + R^+ (X R)] + s(t) [insertions + fixed]
// with jac = 2|t| and s(t) = 2|t| J(t^2).  Integrator and controller: adaptive DOP853 with the
// error norm h e5 / sqrt((e5 + 0.01 e3) N), factor min(10, 0.9 err^-1/8), rejection factor
// max(0.2, .), no step across t = 0, 7th-order dense output at the Gauss nodes, as the reference
// GPU package (gpu/src/rcmps_gpu/optimizer/runtime/ode_backend.py:500-586,
// device_control_backend.py:30-47).  The FSAL stage is evaluated only after acceptance.
//
// Work decomposition: one workgroup per (system, live channel, 16x16 output tile), plus, for the
// kinetic commutator's adjoint system, one E task per tile forming E = X [Q,R] of channel 1, which
// channel 0's insertion reads (so no channel forms two level-1 products).  A stage is
// two levels: level 1 forms the tile of P = R X (forward) or X R (adjoint) and publishes it;
// level 2 forms Q X + X Q^+ (resp. Q^+ X + X Q) while neighbours' P tiles arrive, then adds
// P R^+ (resp. R^+ P), the insertion terms and the next stage input.  Workgroups never meet at
// a grid barrier: each waits on per-tile event flags of the tiles it reads; data written for
// other workgroups uses agent-coherent stores and loads.  The step-end error norm is a
// deterministic all-gather of per-workgroup records, so every workgroup of a system takes the
// same controller decision.  Every wait is bounded by a time budget and watches an abort word.
#include "device_common.h"
#include "split.h"
#include "trajectory.h"

namespace {

template <int NB>
struct Lay {
    static constexpr int Dp = 16 * NB;
    static constexpr int RS = Dp + 1;                  // padded row stride of row blocks
    static constexpr int ROWBLK = 16 * RS;
    static constexpr int COLBLK = Dp * 16;
    static constexpr int BLK = ROWBLK > COLBLK ? ROWBLK : COLBLK;
    static constexpr int XR = 0, XC = BLK, PB = 2 * BLK, K0 = 3 * BLK, K1 = 4 * BLK, K2 = 5 * BLK,
                         K3 = 6 * BLK, K4 = 7 * BLK, RED = 8 * BLK;
    static constexpr int NCPLX = 8 * BLK + 1024;
    static constexpr int NDOUBLE_SCRATCH = 64;
    static constexpr int BYTES = NCPLX * 16 + NDOUBLE_SCRATCH * 8;
};

struct Ctl {
    uint32_t* abort_flag;
    long long wait_ticks;   // longest single wait, in ticks of the constant-rate clock (s_memrealtime)
};

// Wave 0 polls up to 64 flags (lane l: flags[l] >= target); all waves then meet.  Returns false
// when the abort word is set or the wait outlasts its time budget (then the abort word is set).
// The budget is wall-clock time, not a poll count: a workgroup that starts a later round of a
// batched launch early may wait a whole round for partners that are still running.
__device__ bool wait_flags(const Ctl& ctl, uint32_t* const* addr_lane, int n, uint32_t target,
                           int* s_abort) {
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 64) {
        uint32_t* a = lane < n ? addr_lane[lane] : nullptr;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        int bad = 0;
        for (uint32_t polls = 1;; ++polls) {
            // the flag and the abort word are loaded together: one round trip per poll
            const uint32_t fv = a ? ld_flag(a) : 0u;
            const uint32_t ab = ld_flag(ctl.abort_flag);
            const bool ok = a ? (fv >= target) : true;
            if (__all(ok)) break;
            if (ab) { bad = 1; break; }
            // the clock is read every 64th poll only
            if ((polls & 63) == 0 && (long long)(__builtin_amdgcn_s_memrealtime() - t0) > ctl.wait_ticks) {
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0) {
            if (bad) st_flag(ctl.abort_flag, 1u);
            *s_abort = bad;
        }
    }
    __syncthreads();
    return uni(*s_abort) == 0;
}

// Every store another workgroup reads is a device-scope (sc1) store, which completes only once
// it is coherent at device scope; waiting for all of them (vmcnt 0) before the flag store is
// therefore a release for that data, and no L2 writeback is needed (nothing read across
// workgroups is ever written with a non-coherent store).
__device__ __forceinline__ void publish(uint32_t* flag, uint32_t value) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) st_flag(flag, value);
}

// A fragment from a row block M[16][RS]: lane l, k-step ks -> M[l%16][4ks + l/16]
template <int NB>
__device__ __forceinline__ cplx fragA(const cplx* M, int ks, int lane) {
    return M[(lane & 15) * Lay<NB>::RS + 4 * ks + (lane >> 4)];
}
// B fragment from a column block N[Dp][16]: N[4ks + l/16][l%16]
__device__ __forceinline__ cplx fragB(const cplx* N, int ks, int lane) {
    return N[(4 * ks + (lane >> 4)) * 16 + (lane & 15)];
}

// acc += A(row block) * B(col block) over this wave's share of the k-steps, k-tiles in `kmask`
// only (uniform; bit t: k-tile t).  Each wave runs one k-step of every k-tile.
template <int NB>
__device__ __forceinline__ void tile_product(cacc& acc, const cplx* A, const cplx* B, int wave, int lane,
                                             int kmask) {
#pragma unroll
    for (int ks = wave; ks < 4 * NB; ks += 4)
        if ((kmask >> (ks >> 2)) & 1) cmfma(acc, fragA<NB>(A, ks, lane), fragB(B, ks, lane));
}

// Sum the four waves' partial tiles; thread t receives element (t/16, t%16).
__device__ __forceinline__ cplx reduce_tile(const cacc& acc, cplx* red, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 256 + (4 * q + (lane >> 4)) * 16 + (lane & 15)] = cacc_get(acc, q);
    __syncthreads();
    const int t = threadIdx.x;
    cplx s = red[t];
    s = cadd(s, red[256 + t]);
    s = cadd(s, red[512 + t]);
    s = cadd(s, red[768 + t]);
    __syncthreads();
    return s;
}

// Deterministic block sum (same inputs give the same bits in every workgroup).
// a DPP move of a double within each row of 16 lanes (CTRL a compile-time DPP control)
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double x, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                            __builtin_amdgcn_readlane(__double2loint(x), l));
}
// Wave sum in a fixed order: pairs, quads (quad_perm), halves (row_half_mirror) and rows
// (row_mirror) by DPP moves, every lane of a row holding the same row sum (addition commutes),
// then the four row sums in order.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
}

__device__ __forceinline__ double block_sum(double v, double* scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double s = ((scratch[0] + scratch[1]) + scratch[2]) + scratch[3];
    __syncthreads();
    return s;
}
// two block sums with one exchange
__device__ __forceinline__ void block_sum2(double& a, double& b, double* scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        scratch[wave] = a;
        scratch[4 + wave] = b;
    }
    __syncthreads();
    a = ((scratch[0] + scratch[1]) + scratch[2]) + scratch[3];
    b = ((scratch[4] + scratch[5]) + scratch[6]) + scratch[7];
    __syncthreads();
}

// Load a row block (rows 16b.., all Dp columns) of a global Dp x Dp matrix into M[16][RS].
template <int NB, bool COH>
__device__ __forceinline__ void load_rowblk(cplx* M, const double* g, uint32_t gbytes, int b) {
    constexpr int Dp = Lay<NB>::Dp;
    const int t = threadIdx.x;
    cplx v[NB];
    if (COH) {
        __amdgpu_buffer_rsrc_t r = rc_rsrc(g, gbytes);
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            int idx = it * 256 + t, i = idx / Dp, k = idx % Dp;
            v[it] = ld_coh(r, (uint32_t)(((16 * b + i) * Dp + k) * 16));
        }
    } else {
        const cplx* gc = reinterpret_cast<const cplx*>(g);
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            int idx = it * 256 + t, i = idx / Dp, k = idx % Dp;
            v[it] = gc[(16 * b + i) * Dp + k];
        }
    }
#pragma unroll
    for (int it = 0; it < NB; ++it) {
        int idx = it * 256 + t, i = idx / Dp, k = idx % Dp;
        M[i * Lay<NB>::RS + k] = v[it];
    }
}

// Load a column block (all rows, columns 16c..) into N[Dp][16].
template <int NB, bool COH>
__device__ __forceinline__ void load_colblk(cplx* N, const double* g, uint32_t gbytes, int c) {
    constexpr int Dp = Lay<NB>::Dp;
    const int t = threadIdx.x;
    cplx v[NB];
    if (COH) {
        __amdgpu_buffer_rsrc_t r = rc_rsrc(g, gbytes);
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            int idx = it * 256 + t, k = idx / 16, j = idx % 16;
            v[it] = ld_coh(r, (uint32_t)((k * Dp + 16 * c + j) * 16));
        }
    } else {
        const cplx* gc = reinterpret_cast<const cplx*>(g);
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            int idx = it * 256 + t, k = idx / 16, j = idx % 16;
            v[it] = gc[k * Dp + 16 * c + j];
        }
    }
#pragma unroll
    for (int it = 0; it < NB; ++it) {
        int idx = it * 256 + t, k = idx / 16, j = idx % 16;
        N[k * 16 + j] = v[it];
    }
}

// Row block b and column block c of the same matrix, all loads in flight together.
template <int NB>
__device__ __forceinline__ void load_row_col_coh(cplx* M, cplx* N, const double* g, uint32_t gbytes, int b, int c) {
    constexpr int Dp = Lay<NB>::Dp;
    const int t = threadIdx.x;
    __amdgpu_buffer_rsrc_t r = rc_rsrc(g, gbytes);
    cplx vr[NB], vc[NB];
#pragma unroll
    for (int it = 0; it < NB; ++it) {
        int idx = it * 256 + t, i = idx / Dp, k = idx % Dp;
        vr[it] = ld_coh(r, (uint32_t)(((16 * b + i) * Dp + k) * 16));
        int kk = idx / 16, j = idx % 16;
        vc[it] = ld_coh(r, (uint32_t)((kk * Dp + 16 * c + j) * 16));
    }
#pragma unroll
    for (int it = 0; it < NB; ++it) {
        int idx = it * 256 + t, i = idx / Dp, k = idx % Dp;
        M[i * Lay<NB>::RS + k] = vr[it];
        int kk = idx / 16, j = idx % 16;
        N[kk * 16 + j] = vc[it];
    }
}


// Record i of a read-only array through scalar loads, so that every field is known wave-uniform.
template <typename T>
__device__ __forceinline__ T load_scalar(const T* arr, int i) {
    static_assert(sizeof(T) % 8 == 0, "records are copied in 8-byte words");
    const __attribute__((address_space(4))) uint64_t* q =
        (const __attribute__((address_space(4))) uint64_t*)(arr + __builtin_amdgcn_readfirstlane(i));
    T t;
    uint64_t* w = reinterpret_cast<uint64_t*>(&t);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 8); ++k) w[k] = q[k];
    return t;
}

// ---------------------------------------------------------------------------------------
// split precision: LDS layout, panel loads into digit planes, wave-split real products
// ---------------------------------------------------------------------------------------
struct LaySplit {
    static constexpr int P = RC_PANEL_BYTES;
    static constexpr int XR = 0, XC = P, PB = 2 * P, K0 = 3 * P, K1 = 4 * P, K2 = 5 * P, K3 = 6 * P,
                         K4 = 7 * P;
    static constexpr int RED = 8 * P;                  // 4 x 256 doubles
    static constexpr int SCR = RED + 4 * 256 * 8;      // 64 doubles
    static constexpr int BYTES = SCR + 64 * 8;
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ double max4(const cplx* v) {
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) m = fmax(m, fmax(fabs(v[q].re), fabs(v[q].im)));
    return m;
}

// Block maxima of N panels (one value per thread each) -> uniform panel scales S[n].
template <int N>
__device__ __forceinline__ void block_scales(const double* m, int* S, double* scr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int n = 0; n < N; ++n) {
        const double v = wave_max(m[n]);
        if (lane == 0) scr[8 + 4 * n + wave] = v;
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < N; ++n) {
        const double* q = scr + 8 + 4 * n;
        S[n] = uni(rc_split_scale(fmax(fmax(q[0], q[1]), fmax(q[2], q[3]))));
    }
}

// Thread t holds four consecutive k of one row (row t / 16, k = 4 (t % 16) + q) of a row block,
// or of one column (column t / 16, k = 4 (t % 16) + q) of a column block; zero beyond Dp.  The 16
// lanes of one row or column are consecutive, so its largest entry is a 16-lane reduction.
template <int NB, bool COH>
__device__ __forceinline__ void fetch_row4(cplx* v, const double* g, uint32_t gbytes, int b) {
    constexpr int Dp = 16 * NB;
    const int t = threadIdx.x, i = t >> 4, k0 = 4 * (t & 15);
    if (COH) {
        const __amdgpu_buffer_rsrc_t r = rc_rsrc(g, gbytes);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v[q] = k0 < Dp ? ld_coh(r, (uint32_t)(((16 * b + i) * Dp + k0 + q) * 16)) : cmake(0, 0);
    } else {
        const cplx* gc = reinterpret_cast<const cplx*>(g);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = k0 < Dp ? gc[(16 * b + i) * Dp + k0 + q] : cmake(0, 0);
    }
}
template <int NB, bool COH>
__device__ __forceinline__ void fetch_col4(cplx* v, const double* g, uint32_t gbytes, int c) {
    constexpr int Dp = 16 * NB;
    const int t = threadIdx.x, j = t >> 4, k0 = 4 * (t & 15);
    if (COH) {
        const __amdgpu_buffer_rsrc_t r = rc_rsrc(g, gbytes);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v[q] = k0 < Dp ? ld_coh(r, (uint32_t)(((k0 + q) * Dp + 16 * c + j) * 16)) : cmake(0, 0);
    } else {
        const cplx* gc = reinterpret_cast<const cplx*>(g);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = k0 < Dp ? gc[(k0 + q) * Dp + 16 * c + j] : cmake(0, 0);
    }
}
template <int NB>
__device__ __forceinline__ void store_row4(unsigned char* panel, const cplx* v, int S) {
    const int t = threadIdx.x, k0 = 4 * (t & 15);
    if (k0 < 16 * NB) rc_store_digits4(panel, t >> 4, k0, v, ldexp(1.0, S));
}
template <int NB>
__device__ __forceinline__ void store_col4(unsigned char* panel, const cplx* v, int S) {
    const int t = threadIdx.x, k0 = 4 * (t & 15);
    if (k0 < 16 * NB) rc_store_digits4(panel, t >> 4, k0, v, ldexp(1.0, S));
}

// biased FP64 exponent of the largest |re|, |im| of v[0..3] (0 for zeros and subnormals)
__device__ __forceinline__ int max_exp4(const cplx* v) {
    int e = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        e = max(e, (int)((uint32_t)__double2hiint(v[q].re) >> 20) & 0x7ff);
        e = max(e, (int)((uint32_t)__double2hiint(v[q].im) >> 20) & 0x7ff);
    }
    return e;
}
// scale of this thread's row (or column) of a row (column) block: the maximum over its 16 lanes.
// The lane holding k = 0 writes it to sg[row or column].
__device__ __forceinline__ int line_scale(const cplx* v, int* sg) {
    int e = max_exp4(v);
    // 16-lane maximum by DPP moves within each row of 16 lanes: quad_perm [1,0,3,2] and [2,3,0,1]
    // give the quad's maximum, row_half_mirror joins quads 0-1 and 2-3, row_mirror the two halves
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0xB1, 0xF, 0xF, false));
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x4E, 0xF, 0xF, false));
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x141, 0xF, 0xF, false));
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x140, 0xF, 0xF, false));
    const int S = rc_split_scale_e(e);
    if ((threadIdx.x & 15) == 0) sg[threadIdx.x >> 4] = S;
    return S;
}

// one constant panel (row block b or column block c of a fixed matrix) at kernel start
template <int NB>
__device__ __forceinline__ int const_panel(unsigned char* panel, const double* g, bool row, int blk, double* scr) {
    cplx v[4];
    if (row) fetch_row4<NB, false>(v, g, 0, blk);
    else fetch_col4<NB, false>(v, g, 0, blk);
    const double m = max4(v);
    int S;
    block_scales<1>(&m, &S, scr);
    if (row) store_row4<NB>(panel, v, S);
    else store_col4<NB>(panel, v, S);
    __syncthreads();
    return S;
}

// Waves 0..3 hold the real products re.re, im.im, re.im, im.re of the same complex tile product;
// thread t receives element (t/16, t%16) of the complex result.
__device__ __forceinline__ cplx reduce_split(const double* part, double* red, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 256 + (4 * (lane >> 4) + q) * 16 + (lane & 15)] = part[q];
    __syncthreads();
    const int t = threadIdx.x;
    const cplx s = cmake(red[t] - red[256 + t], red[512 + t] + red[768 + t]);
    __syncthreads();
    return s;
}

// DOP853 combinations sum_j c_j k_j with the tableau as compile-time constants: zero terms
// vanish and every index of the stage-derivative window k is static, so k stays in registers.
// k_j sits at k[OFF + j].  The sum runs over j in increasing order from zero, as a running
// accumulation would.
enum RcRow { RC_ROW_A, RC_ROW_B, RC_ROW_E5, RC_ROW_E3, RC_ROW_D };

template <RcRow W, int S>
__device__ __forceinline__ constexpr double rc_coef(int j) {
    if constexpr (W == RC_ROW_A) return RC_A[S][j];
    if constexpr (W == RC_ROW_B) return j < RC_NSTAGES ? RC_B[j] : 0.0;
    if constexpr (W == RC_ROW_E5) return j < RC_NSTAGES ? RC_E5[j] : 0.0;
    if constexpr (W == RC_ROW_E3) return j < RC_NSTAGES ? RC_E3[j] : 0.0;
    return RC_D[S][j];
}

template <RcRow W, int S, int N, int OFF, int J = 0>
__device__ __forceinline__ cplx rc_comb(const cplx* k, cplx acc) {
    if constexpr (J == N) {
        return acc;
    } else {
        constexpr double c = rc_coef<W, S>(J);
        if constexpr (c != 0.0) acc = cadd(acc, cscale(c, k[OFF + J]));
        return rc_comb<W, S, N, OFF, J + 1>(k, acc);
    }
}

// sum_{j < s} A[s][j] k_j for a uniform stage index s in 1..15, taken right after k_{s-1}
// entered the window (k_j at k[16 - s + j])
template <int S = 1>
__device__ __forceinline__ cplx rc_stage_comb(int s, const cplx* k) {
    if constexpr (S == RC_NSTAGES_EXT - 1) {
        return rc_comb<RC_ROW_A, S, S, RC_NSTAGES_EXT - S>(k, cmake(0, 0));
    } else {
        if (s == S) return rc_comb<RC_ROW_A, S, S, RC_NSTAGES_EXT - S>(k, cmake(0, 0));
        return rc_stage_comb<S + 1>(s, k);
    }
}

}  // namespace

template <int NB, bool SPLIT>
__device__ void run_task(const TrajParams& p, int task, unsigned char* smem_raw) {
    using L = Lay<NB>;
    constexpr int Dp = L::Dp, NT = NB * NB;
    cplx* sm = reinterpret_cast<cplx*>(smem_raw);
    cplx* Xr = sm + L::XR;
    cplx* Xc = sm + L::XC;
    cplx* Pb = sm + L::PB;
    cplx* Kl1 = sm + L::K0;   // level-1 constant: R row block (fwd) / R col block (adj)
    cplx* Kq = sm + L::K1;    // Q row block (fwd) / Q^+ row block (adj)
    cplx* Kx = sm + L::K2;    // Q^+ col block (fwd) / Q col block (adj)
    cplx* Kp = sm + L::K3;    // R^+ col block (fwd) / R^+ row block (adj)
    cplx* Kc = sm + L::K4;    // C col block (adj channel with E)
    cplx* red = sm + L::RED;
    // split precision: digit planes of the same operands (FP64 blocks above unused)
    unsigned char* pXr = smem_raw + LaySplit::XR;
    unsigned char* pXc = smem_raw + LaySplit::XC;
    unsigned char* pPb = smem_raw + LaySplit::PB;
    unsigned char* pK0 = smem_raw + LaySplit::K0;
    unsigned char* pK1 = smem_raw + LaySplit::K1;
    unsigned char* pK2 = smem_raw + LaySplit::K2;
    unsigned char* pK3 = smem_raw + LaySplit::K3;
    unsigned char* pK4 = smem_raw + LaySplit::K4;
    double* redd = reinterpret_cast<double*>(smem_raw + LaySplit::RED);
    double* scratch = SPLIT ? reinterpret_cast<double*>(smem_raw + LaySplit::SCR)
                            : reinterpret_cast<double*>(sm + L::NCPLX);
    __shared__ int s_abort;
    __shared__ uint32_t* s_addr[64];
    __shared__ double s_tot[16];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int spa = (wave == 1 || wave == 3), spb = (wave == 1 || wave == 2);   // this wave's real product
    const int ei = tid >> 4, ej = tid & 15;
    const TrajTask tk = load_scalar<TrajTask>(p.tasks, task);
    const TrajSys S = load_scalar<TrajSys>(p.sys, tk.sys);
    const int b = tk.b, c = tk.c, ch = tk.ch, adj = tk.adj;
    // k-tile masks of the tile products (uniform)
    const int m1 = tk.kmask & 15, m2a = (tk.kmask >> 4) & 15, m2b = (tk.kmask >> 8) & 15,
              m3 = (tk.kmask >> 12) & 15, mE = (tk.kmask >> 16) & 15;
    const int row = 16 * b + ei, col = 16 * c + ej;
    const bool e_only = tk.e_only != 0;                  // uniform over the workgroup
    const bool real_el = !e_only && row < S.D && col < S.D;
    const Ctl ctl{p.abort_flag, p.wait_ticks};
    const uint32_t mat_bytes = (uint32_t)(Dp * Dp * 16);
    const uint32_t stage_stride = (uint32_t)S.nlive * mat_bytes;
    const uint32_t ybytes = RC_NSTAGE_BUF * stage_stride;
    const __amdgpu_buffer_rsrc_t rY = rc_rsrc(S.Y, ybytes);
    const __amdgpu_buffer_rsrc_t rP = rc_rsrc(S.P, ybytes);
    const __amdgpu_buffer_rsrc_t rE = rc_rsrc(S.E, S.E ? RC_NSTAGE_BUF * mat_bytes : 0);
    const uint32_t el_off = (uint32_t)((row * Dp + col) * 16);
    const uint32_t el_off_t = (uint32_t)((col * Dp + row) * 16);   // transposed position

    // constant blocks
    int sK0 = 0, sK1 = 0, sK2 = 0, sK3 = 0, sK4 = 0;   // split-precision panel scales
    // per-row / per-column scales of the stage input's row block [0, 16) and column block
    // [16, 32), and of the P block (split precision)
    int* sgX = reinterpret_cast<int*>(scratch + 16);
    int* sgP = sgX + 32;
    if constexpr (SPLIT) {
        // zero every digit plane once: k beyond Dp stays zero
        for (int i = tid; i < 8 * RC_PANEL_BYTES / 16; i += 256) reinterpret_cast<v4i*>(smem_raw)[i] = v4i{0, 0, 0, 0};
        __syncthreads();
        if (!adj) {
            sK0 = const_panel<NB>(pK0, S.R, true, b, scratch);
            sK1 = const_panel<NB>(pK1, S.Q, true, b, scratch);
            sK2 = const_panel<NB>(pK2, S.Qh, false, c, scratch);
            sK3 = const_panel<NB>(pK3, S.Rh, false, c, scratch);
        } else {
            sK0 = const_panel<NB>(pK0, S.R, false, c, scratch);
            sK1 = const_panel<NB>(pK1, S.Qh, true, b, scratch);
            sK2 = const_panel<NB>(pK2, S.Q, false, c, scratch);
            sK3 = const_panel<NB>(pK3, S.Rh, true, b, scratch);
            if (tk.has_E) sK4 = const_panel<NB>(pK4, S.C, false, c, scratch);
        }
    } else if (!adj) {
        load_rowblk<NB, false>(Kl1, S.R, 0, b);
        load_rowblk<NB, false>(Kq, S.Q, 0, b);
        load_colblk<NB, false>(Kx, S.Qh, 0, c);
        load_colblk<NB, false>(Kp, S.Rh, 0, c);
    } else {
        load_colblk<NB, false>(Kl1, S.R, 0, c);
        load_rowblk<NB, false>(Kq, S.Qh, 0, b);
        load_colblk<NB, false>(Kx, S.Q, 0, c);
        load_rowblk<NB, false>(Kp, S.Rh, 0, b);
        if (tk.has_E) load_colblk<NB, false>(Kc, S.C, 0, c);
    }
    cplx fixed_e = cmake(0, 0), trop_e = cmake(0, 0);
    if (tk.fixed) fixed_e = reinterpret_cast<const cplx*>(tk.fixed)[row * Dp + col];
    if (tk.trace_kind == 1) trop_e = reinterpret_cast<const cplx*>(tk.trace_op)[row * Dp + col];
    __syncthreads();

    uint32_t* flagY = S.flagY + ch * NT;
    uint32_t* flagP = S.flagP + ch * NT;
    const int my_tile = b * NB + c;

    // ------------------------------------------------------------------------------------
    // one right-hand-side evaluation on the stage input published in slot si with event ev
    // ------------------------------------------------------------------------------------
    // phase timers (constant-rate clock ticks): wait Y, load X, level 1 + publish P, level 2a,
    // wait P, load P/insertions, level 2b + epilogue, integrator bookkeeping, step end (error
    // norm, record, all-gather), total, publish Y, stage entry
    // The timers run only when the caller asks for them: each reading of the clock waits for the
    // scalar memory counter, which the untimed kernel does not pay.
    const bool timed = uni(p.timers) != 0;
    uint64_t tacc[RC_NTIMERS];
#pragma unroll
    for (int q = 0; q < RC_NTIMERS; ++q) tacc[q] = 0;
    uint64_t tmark = timed ? __builtin_amdgcn_s_memrealtime() : 0;
    const uint64_t tstart = tmark;
    auto lap = [&](int q) {
        if (timed) {
            uint64_t now = __builtin_amdgcn_s_memrealtime();
            tacc[q] += now - tmark;
            tmark = now;
        }
    };
    auto stage = [&](int si, uint32_t ev, double ts, double src, cplx& k_out, double& tr_out) -> bool {
        const double jac = 2.0 * fabs(ts);
        // wait for own channel's column block c and row block b
        if (tid < 64) {
            int n = 0;
            if (lane < NB) s_addr[lane] = flagY + lane * NB + c;                    // (lane, c)
            if (lane < NB) s_addr[NB + lane] = flagY + b * NB + lane;               // (b, lane)
            (void)n;
        }
        __syncthreads();
        lap(11);
        if (!wait_flags(ctl, s_addr, 2 * NB, ev, &s_abort)) return false;
        lap(0);
        const double* Ych = S.Y + ((size_t)si * S.nlive + ch) * Dp * Dp * 2;
        cplx xel = cmake(0, 0);   // own element of the stage input (split path)
        if constexpr (SPLIT) {
            cplx vr[4], vc[4];
            fetch_row4<NB, true>(vr, Ych, mat_bytes, b);
            fetch_col4<NB, true>(vc, Ych, mat_bytes, c);
            if (tk.trace_kind == 1) xel = ld_coh(rY, (uint32_t)si * stage_stride + (uint32_t)ch * mat_bytes + el_off);
            store_row4<NB>(pXr, vr, line_scale(vr, sgX));
            store_col4<NB>(pXc, vc, line_scale(vc, sgX + 16));
        } else {
            load_row_col_coh<NB>(Xr, Xc, Ych, mat_bytes, b, c);
        }
        __syncthreads();
        // per-element product scales: this lane's output rows 4 (lane/16) + q and column lane%16
        int sXr[4], sXc = 0;
        if constexpr (SPLIT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sXr[q] = sgX[4 * (lane >> 4) + q];
            sXc = sgX[16 + (lane & 15)];
        }
        lap(1);
        if (e_only) {
            // E = X C of channel ch: its row block b times C's column block c
            cplx Ee;
            if constexpr (SPLIT) {
                double pe[4] = {0.0, 0.0, 0.0, 0.0};
                const int s2[4] = {-(sXr[0] + sK4), -(sXr[1] + sK4), -(sXr[2] + sK4), -(sXr[3] + sK4)};
                if (mE) rc_split_product(pe, pXr, spa, pK4, spb, s2, lane);
                Ee = reduce_split(pe, redd, wave, lane);
            } else {
                cacc ae;
                cacc_zero(ae);
                tile_product<NB>(ae, Xr, Kc, wave, lane, mE);
                Ee = reduce_tile(ae, red, wave, lane);
            }
            st_coh(rE, (uint32_t)si * mat_bytes + el_off, Ee);
            publish(S.flagE + my_tile, ev);
            k_out = cmake(0, 0);
            tr_out = 0.0;
            return true;
        }
        // level 1
        cacc acc;
        double part[4] = {0.0, 0.0, 0.0, 0.0};
        cplx Pe;
        if constexpr (SPLIT) {
            if (!adj) {
                const int s2[4] = {-(sK0 + sXc), -(sK0 + sXc), -(sK0 + sXc), -(sK0 + sXc)};
                if (m1) rc_split_product(part, pK0, spa, pXc, spb, s2, lane);
            } else {
                const int s2[4] = {-(sXr[0] + sK0), -(sXr[1] + sK0), -(sXr[2] + sK0), -(sXr[3] + sK0)};
                if (m1) rc_split_product(part, pXr, spa, pK0, spb, s2, lane);
            }
            Pe = reduce_split(part, redd, wave, lane);
        } else {
            cacc_zero(acc);
            if (!adj) tile_product<NB>(acc, Kl1, Xc, wave, lane, m1);
            else tile_product<NB>(acc, Xr, Kl1, wave, lane, m1);
            Pe = reduce_tile(acc, red, wave, lane);
        }
        const uint32_t poff = (uint32_t)si * stage_stride + (uint32_t)ch * mat_bytes;
        st_coh(rP, poff + el_off, Pe);
        double trv = 0.0;
        if (tk.trace_kind == 2 && b == c && ei == ej && real_el) trv = Pe.re;
        if (tk.trace_kind == 1 && real_el) {
            const cplx x = SPLIT ? xel : Xr[ei * L::RS + col];
            trv = x.re * trop_e.re + x.im * trop_e.im;   // Re(x conj(o))
        }
        publish(flagP + my_tile, ev);
        lap(2);
        // level 2a
        if constexpr (SPLIT) {
            part[0] = part[1] = part[2] = part[3] = 0.0;
            const int s2a[4] = {-(sK1 + sXc), -(sK1 + sXc), -(sK1 + sXc), -(sK1 + sXc)};
            const int s2b[4] = {-(sXr[0] + sK2), -(sXr[1] + sK2), -(sXr[2] + sK2), -(sXr[3] + sK2)};
            if (m2a) rc_split_product(part, pK1, spa, pXc, spb, s2a, lane);
            if (m2b) rc_split_product(part, pXr, spa, pK2, spb, s2b, lane);
        } else {
            cacc_zero(acc);
            tile_product<NB>(acc, Kq, Xc, wave, lane, m2a);
            tile_product<NB>(acc, Xr, Kx, wave, lane, m2b);
        }
        __syncthreads();
        lap(3);
        // wait for the P tiles of this channel and the insertion source tiles
        if (tid < 64) {
            if (lane < NB) s_addr[lane] = adj ? (flagP + lane * NB + c) : (flagP + b * NB + lane);
            if (lane == 0 && tk.edge_src >= 0) {
                uint32_t* fs = tk.edge_from_E ? S.flagE : (S.flagP + tk.edge_src * NT);
                s_addr[NB] = fs + b * NB + c;
                s_addr[NB + 1] = fs + c * NB + b;
            }
        }
        __syncthreads();
        const int nw = NB + (tk.edge_src >= 0 ? 2 : 0);
        if (!wait_flags(ctl, s_addr, nw, ev, &s_abort)) return false;
        lap(4);
        const double* Pch = S.P + ((size_t)si * S.nlive + ch) * Dp * Dp * 2;
        cplx vp[4];
        if constexpr (SPLIT) {
            if (!adj) fetch_row4<NB, true>(vp, Pch, mat_bytes, b);
            else fetch_col4<NB, true>(vp, Pch, mat_bytes, c);
        } else {
            if (!adj) load_rowblk<NB, true>(Pb, Pch, mat_bytes, b);
            else load_colblk<NB, true>(Pb, Pch, mat_bytes, c);
        }
        cplx ins = cmake(0, 0);
        if (tk.edge_src >= 0) {
            cplx s1, s2;
            if (tk.edge_from_E) {
                s1 = ld_coh(rE, (uint32_t)si * mat_bytes + el_off);
                s2 = ld_coh(rE, (uint32_t)si * mat_bytes + el_off_t);
            } else {
                uint32_t so = (uint32_t)si * stage_stride + (uint32_t)tk.edge_src * mat_bytes;
                s1 = ld_coh(rP, so + el_off);
                s2 = ld_coh(rP, so + el_off_t);
            }
            ins = cscale(tk.edge_coef, cadd(s1, cconj(s2)));
        }
        if constexpr (SPLIT) {
            if (!adj) store_row4<NB>(pPb, vp, line_scale(vp, sgP));
            else store_col4<NB>(pPb, vp, line_scale(vp, sgP));
        }
        __syncthreads();
        lap(5);
        cplx Le;
        if constexpr (SPLIT) {
            if (!adj) {
                const int s2[4] = {-(sgP[4 * (lane >> 4)] + sK3), -(sgP[4 * (lane >> 4) + 1] + sK3),
                                   -(sgP[4 * (lane >> 4) + 2] + sK3), -(sgP[4 * (lane >> 4) + 3] + sK3)};
                if (m3) rc_split_product(part, pPb, spa, pK3, spb, s2, lane);
            } else {
                const int sc = -(sK3 + sgP[lane & 15]);
                const int s2[4] = {sc, sc, sc, sc};
                if (m3) rc_split_product(part, pK3, spa, pPb, spb, s2, lane);
            }
            Le = reduce_split(part, redd, wave, lane);
        } else {
            if (!adj) tile_product<NB>(acc, Pb, Kp, wave, lane, m3);
            else tile_product<NB>(acc, Kp, Pb, wave, lane, m3);
            Le = reduce_tile(acc, red, wave, lane);
        }
        cplx k = cadd(cscale(jac, Le), cscale(src, cadd(ins, fixed_e)));
        if (!real_el) k = cmake(0, 0);
        k_out = k;
        if (tk.trace_kind) tr_out = tk.trace_coef * src * block_sum(trv, scratch);
        else tr_out = 0.0;
        lap(6);
        return true;
    };

    auto publish_Y = [&](int si, uint32_t ev, cplx v) {
        lap(7);
        if (!e_only) {   // an E task reads channel ch's stage input; the channel's own task writes it
            st_coh(rY, (uint32_t)si * stage_stride + (uint32_t)ch * mat_bytes + el_off, v);
            publish(flagY + my_tile, ev);
        }
        lap(10);
    };

    // ------------------------------------------------------------------------------------
    // integrator: a state machine over stage slots with a single stage call site.  The stage
    // derivatives of the attempt are kept per element in a window kst that shifts by one as each
    // arrives (the newest at kst[15]); every combination (next stage input, solution, error
    // estimators, dense output) is a sum with the tableau's constant coefficients, selected by
    // the uniform slot, so no array is indexed at run time.
    // ------------------------------------------------------------------------------------
    const double a = p.a;
    cplx kst[RC_NSTAGES_EXT];
#pragma unroll
    for (int q = 0; q < RC_NSTAGES_EXT; ++q) kst[q] = cmake(0, 0);
    cplx y = cmake(0, 0), ynew = cmake(0, 0), k0 = cmake(0, 0), k12 = cmake(0, 0);
    __shared__ double s_trp[RC_NSTAGES_EXT];
    __shared__ double s_src[RC_NSTAGES_EXT];
    __shared__ double s_cheb[RC_SOURCE_SEGMENTS * (RC_SOURCE_DEGREE + 1)];
    for (int i = tid; i < RC_SOURCE_SEGMENTS * (RC_SOURCE_DEGREE + 1); i += 256) s_cheb[i] = p.cheb[i];
    __shared__ double s_cnode[RC_NSTAGES_EXT];
    if (tid < RC_NSTAGES_EXT) s_cnode[tid] = RC_C[tid];
    __syncthreads();

    double ysc = 0.0, ysc_new = 0.0;
    double t = -a, h = fmin(p.h0, 2.0 * a), t_new = t, endpoint = 0.0, factor = 1.0;
    uint32_t ev = 0;
    long long attempts = 0, accepted = 0, rejected = 0, dense = 0;
    int next_point = 0, rej = 0, status = RC_ST_OK;
    bool alive = true;
    const __amdgpu_buffer_rsrc_t rS = rc_rsrc(S.samples, (uint32_t)((size_t)p.npoints * S.nlive * mat_bytes));
    const int my_rank = task - S.first_task;

    int slot = 0;
    cplx Yin = y;
    double ts = t;
    while (true) {
        ++ev;
        publish_Y(slot, ev, Yin);
        cplx kn;
        double trn;
        const double src = slot == 0 ? rc_source(ts, s_cheb) : unid(s_src[slot]);
        if (!(alive = stage(slot, ev, ts, src, kn, trn))) break;
        if (tid == 0) s_trp[slot] = trn;
        int next = -1;   // -1: start a new step attempt
#pragma unroll
        for (int q = 0; q < RC_NSTAGES_EXT - 1; ++q) kst[q] = kst[q + 1];
        kst[RC_NSTAGES_EXT - 1] = kn;
        if (slot == 0) {
            k0 = kn;
        } else if (slot < RC_NSTAGES - 1) {
            next = slot + 1;
        } else if (slot == RC_NSTAGES - 1) {
            lap(7);
            // k_j at kst[4 + j] after k_11
            const cplx accB = rc_comb<RC_ROW_B, 0, RC_NSTAGES, 4>(kst, cmake(0, 0));
            const cplx acc5 = rc_comb<RC_ROW_E5, 0, RC_NSTAGES, 4>(kst, cmake(0, 0));
            const cplx acc3 = rc_comb<RC_ROW_E3, 0, RC_NSTAGES, 4>(kst, cmake(0, 0));
            ynew = cadd(y, cscale(h, accB));
            double e5p = 0.0, e3p = 0.0;
            if (real_el) {
                double sc = p.atol + p.rtol * fmax(cabsd(y), cabsd(ynew));
                e5p = cabs2(cmake(acc5.re / sc, acc5.im / sc));
                e3p = cabs2(cmake(acc3.re / sc, acc3.im / sc));
            }
            block_sum2(e5p, e3p, scratch);
            // record for the all-gather: e5, e3, trace partials of k_0..k_11
            const uint32_t par = (uint32_t)(attempts & 1);
            if (tid < RC_REC) {
                double v = 0.0;
                if (tid == 0) v = e5p;
                else if (tid == 1) v = e3p;
                else if (tid < 2 + RC_NSTAGES) v = s_trp[tid - 2];
                st_coh_f64(S.rec + ((size_t)par * S.ntasks + my_rank) * RC_REC + tid, v);
            }
            lap(8);
            publish(S.flagR + my_rank, (uint32_t)attempts);
            if (tid < 64) s_addr[lane] = S.flagR + (lane < S.ntasks ? lane : 0);
            __syncthreads();
            if (!(alive = wait_flags(ctl, s_addr, S.ntasks, (uint32_t)attempts, &s_abort))) break;
            if (tid < 64) {
                // every record in flight at once, then one deterministic reduction per entry
                const __amdgpu_buffer_rsrc_t rR = rc_rsrc(S.rec, (uint32_t)(2 * S.ntasks * RC_REC * 8));
                const uint32_t base = (uint32_t)(((size_t)par * S.ntasks + lane) * RC_REC * 8);
                cplx v[(2 + RC_NSTAGES) / 2];
#pragma unroll
                for (int q = 0; q < (2 + RC_NSTAGES) / 2; ++q)
                    v[q] = lane < S.ntasks ? ld_coh(rR, base + 16 * q) : cmake(0, 0);
#pragma unroll
                for (int q = 0; q < (2 + RC_NSTAGES) / 2; ++q) {
                    double x = v[q].re, z = v[q].im;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) {
                        x += __shfl_xor(x, off);
                        z += __shfl_xor(z, off);
                    }
                    if (lane == 0) {
                        s_tot[2 * q] = x;
                        s_tot[2 * q + 1] = z;
                    }
                }
            }
            __syncthreads();
            lap(8);
            double e5 = unid(s_tot[0]), e3 = unid(s_tot[1]);
            ysc_new = ysc;
            if (S.has_scalar) {
                double sb = 0.0, s5 = 0.0, s3 = 0.0;
#pragma unroll
                for (int j = 0; j < RC_NSTAGES; ++j) {
                    const double kj = unid(s_tot[2 + j]);
                    sb += RC_B[j] * kj;
                    s5 += RC_E5[j] * kj;
                    s3 += RC_E3[j] * kj;
                }
                ysc_new = ysc + h * sb;
                double sc = p.atol + p.rtol * fmax(fabs(ysc), fabs(ysc_new));
                e5 += (s5 / sc) * (s5 / sc);
                e3 += (s3 / sc) * (s3 / sc);
            }
            __syncthreads();
            const double den = e5 + 0.01 * e3;
            const double error = den > 0.0 ? h * e5 / sqrt(den * S.n_elems) : 0.0;
            if (!isfinite(error)) { status = RC_ST_NONFINITE; break; }
            factor = error > 0.0 ? fmin(10.0, 0.9 * pow(error, -1.0 / 8.0)) : 10.0;
            if (error > 1.0) {
                ++rejected;
                rej = 1;
                h *= fmax(0.2, factor);
                next = -1;
            } else {
                ++accepted;
                t_new = (h == endpoint - t) ? endpoint : t + h;
                next = RC_NSTAGES;   // FSAL stage
            }
        } else if (slot == RC_NSTAGES) {
            k12 = kn;
            if (next_point < p.npoints && ld_const(p.points, next_point) <= t_new) {
                ++dense;
                next = RC_NSTAGES + 1;
            } else {
                next = -2;   // finish the accepted step
            }
        } else if (slot < RC_NSTAGES_EXT - 1) {
            next = slot + 1;
        } else {
            const cplx d0 = csub(ynew, y);
            const cplx d1 = csub(cscale(h, k0), d0);
            const cplx d2 = csub(cscale(2.0, d0), cscale(h, cadd(k12, k0)));
            // k_j at kst[j] after k_15
            cplx dd[4];
            dd[0] = cscale(h, rc_comb<RC_ROW_D, 0, RC_NSTAGES_EXT, 0>(kst, cmake(0, 0)));
            dd[1] = cscale(h, rc_comb<RC_ROW_D, 1, RC_NSTAGES_EXT, 0>(kst, cmake(0, 0)));
            dd[2] = cscale(h, rc_comb<RC_ROW_D, 2, RC_NSTAGES_EXT, 0>(kst, cmake(0, 0)));
            dd[3] = cscale(h, rc_comb<RC_ROW_D, 3, RC_NSTAGES_EXT, 0>(kst, cmake(0, 0)));
            while (next_point < p.npoints && ld_const(p.points, next_point) <= t_new) {
                const double pt = ld_const(p.points, next_point);
                cplx v;
                if (pt == t_new) {
                    v = ynew;
                } else {
                    const double x = (pt - t) / h, xm = 1.0 - x;
                    cplx r = cscale(x, dd[3]);
                    r = cscale(xm, cadd(r, dd[2]));
                    r = cscale(x, cadd(r, dd[1]));
                    r = cscale(xm, cadd(r, dd[0]));
                    r = cscale(x, cadd(r, d2));
                    r = cscale(xm, cadd(r, d1));
                    r = cscale(x, cadd(r, d0));
                    v = cadd(r, y);
                }
                if (!e_only) st_coh(rS, (uint32_t)next_point * stage_stride + (uint32_t)ch * mat_bytes + el_off, v);
                ++next_point;
            }
            next = -2;
        }
        if (next == -2) {
            // accepted step complete: FSAL and state update
            y = ynew;
            k0 = k12;
            if (tid == 0) s_trp[0] = s_trp[RC_NSTAGES];
            t = t_new;
            ysc = ysc_new;
            h *= rej ? fmin(1.0, factor) : factor;
            rej = 0;
            next = -1;
        }
        if (next == -1) {
            // start an attempt from (t, y, k0) with step h
            if (!(t < a)) break;
            endpoint = t < 0.0 ? 0.0 : a;
            h = fmin(h, endpoint - t);
            if (!(h > 10.0 * fabs(nextafter(t, INFINITY) - t))) { status = RC_ST_STEP_UNDERFLOW; break; }
            if (attempts >= p.max_steps) { status = RC_ST_MAX_STEPS; break; }
            ++attempts;
            kst[RC_NSTAGES_EXT - 1] = k0;   // k_0 of the attempt enters the window
            // source values at every stage time of this attempt (stages 1..15), one lane each
            if (tid >= 1 && tid < RC_NSTAGES_EXT) s_src[tid] = rc_source(t + s_cnode[tid] * h, s_cheb);
            __syncthreads();
            next = 1;
        }
        slot = uni(next);
        next_point = uni(next_point);
        t = unid(t);
        h = unid(h);
        if (slot == RC_NSTAGES) {
            Yin = ynew;
            ts = t + h;
        } else {
            Yin = cadd(y, cscale(h, rc_stage_comb(slot, kst)));
            ts = t + RC_C[slot] * h;
        }
    }
    if (!alive) status = RC_ST_WAIT_LIMIT;
    if (status == RC_ST_OK && next_point != p.npoints) status = RC_ST_MISSING_SAMPLES;
    __builtin_amdgcn_s_waitcnt(0);
    if (my_rank == 0 && tid == 0) {
        S.out[RC_OUT_STATUS] = status;
        S.out[RC_OUT_ATTEMPTS] = (double)attempts;
        S.out[RC_OUT_ACCEPTED] = (double)accepted;
        S.out[RC_OUT_REJECTED] = (double)rejected;
        S.out[RC_OUT_DENSE] = (double)dense;
        S.out[RC_OUT_RHS] = (double)ev;
        S.out[RC_OUT_SCALAR] = ysc;
        S.out[RC_OUT_T] = t;
        S.out[RC_OUT_NEXTPOINT] = next_point;
        lap(7);
        if (timed) tacc[9] = __builtin_amdgcn_s_memrealtime() - tstart;
#pragma unroll
        for (int q = 0; q < RC_NTIMERS; ++q) S.out[RC_OUT_TIMERS + q] = (double)tacc[q];
    }
}

template <int NB, bool SPLIT>
__global__ void __launch_bounds__(256, 1) rc_trajectory(const TrajParams* __restrict__ pp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const TrajParams& p = *pp;
    const __attribute__((address_space(4))) int* sched = (const __attribute__((address_space(4))) int*)p.schedule;
    for (int r = 0; r < p.nrounds; ++r) {
        const int task = __builtin_amdgcn_readfirstlane(sched[r * (int)gridDim.x + (int)blockIdx.x]);
        if (task < 0) continue;
        run_task<NB, SPLIT>(p, task, smem_raw);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// host launcher
// ---------------------------------------------------------------------------------------
int rc_trajectory_smem_bytes(int nb, int split) {
    if (nb < 1 || nb > 4) return -1;
    if (split) return LaySplit::BYTES;
    switch (nb) {
        case 1: return Lay<1>::BYTES;
        case 2: return Lay<2>::BYTES;
        case 3: return Lay<3>::BYTES;
        default: return Lay<4>::BYTES;
    }
}

static const void* trajectory_fn(int nb, int split) {
    switch (nb * 2 + (split ? 1 : 0)) {
        case 2: return (const void*)rc_trajectory<1, false>;
        case 3: return (const void*)rc_trajectory<1, true>;
        case 4: return (const void*)rc_trajectory<2, false>;
        case 5: return (const void*)rc_trajectory<2, true>;
        case 6: return (const void*)rc_trajectory<3, false>;
        case 7: return (const void*)rc_trajectory<3, true>;
        case 8: return (const void*)rc_trajectory<4, false>;
        case 9: return (const void*)rc_trajectory<4, true>;
        default: return nullptr;
    }
}

// workgroups of the trajectory kernel that can be resident at once on the current device
hipError_t rc_trajectory_capacity(int nb, int split, int* blocks) {
    const void* fn = trajectory_fn(nb, split);
    if (!fn) return hipErrorInvalidValue;
    const int bytes = rc_trajectory_smem_bytes(nb, split);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    int per_cu = 0, dev = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, bytes)) != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    *blocks = per_cu * cus;
    return hipSuccess;
}

hipError_t rc_launch_trajectory(int nb, int split, int nblocks, const TrajParams* dparams, hipStream_t stream) {
    void* args[] = {(void*)&dparams};
    const void* fn = trajectory_fn(nb, split);
    const int bytes = rc_trajectory_smem_bytes(nb, split);
    if (!fn) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(256), args, (unsigned)bytes, stream);
}
