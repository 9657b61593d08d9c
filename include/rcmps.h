/*
 * C ABI of librcmps_hip.so: the phi^4 energy-and-gradient path of RcmpsCore on one MI355X.
 *
 * Matrices crossing the boundary are DEVICE pointers to row-major D x D complex128 arrays
 * (interleaved re, im doubles; element (i, j) at offset 2*(i*D + j)).  The caller owns them.
 * Scalars come back through host pointers.  Every entry point returns RCMPS_OK (0) or a
 * status code; rcmps_status_string() names it.  A context owns one HIP stream and its
 * workspaces; calls on one context are serialised by the caller, distinct contexts are
 * independent.
 *
 * Reference interfaces replaced (AntoineTilloy/RcmpsCore.jl):
 *   rcmps_state_create        State(R, K) with its stationary state      src/state.jl:40-53,
 *                             (CPTP Arnoldi / Poisson solve of rho0)      src/transfer.jl:74-115;
 *                             rcmps_gpu numerics.stationary_state          gpu/src/rcmps_gpu/optimizer/
 *                                                                          runtime/numerics.py:74-100
 *   rcmps_fg_phi4             fg_phi4(s, g) -> (energy, gradient)         src/observables.jl:226-257;
 *                             rcmps_gpu objective contract                 gpu/src/rcmps_gpu/optimizer/
 *                                                                          runtime/sweep_driver.py:78-118
 *   rcmps_boundary_environment  boundary_environment(s, a)                src/transfer.jl:122-153
 *   rcmps_trajectory_samples  the dense-output samples of one trajectory  src/observables.jl:77-96
 *                             (forward/adjoint trajectory solutions)
 *   rcmps_excitation_create   ExcitationContext(s, p)                     src/excitation_environments.jl:18-40
 *   rcmps_excitation_poisson  excitation_poisson, action_dual_poisson     src/excitation_environments.jl:75-85,
 *                                                                          src/excitation_action.jl:85-94
 *   rcmps_excitation_element  excitation_element(ExcitationPair(ctx,U,W))  src/excitations.jl:98-113
 *   rcmps_excitation_action   excitation_hamiltonian_action(ctx, W)       src/excitation_action.jl:194-228
 *                             (excitation_kinetic_action_sum :135-179, excitation_observable_action :103-126);
 *                             rcmps_gpu HamiltonianAction                  gpu/src/rcmps_gpu/excitations/
 *                                                                          hamiltonian.py:21-72
 */
#ifndef RCMPS_H
#define RCMPS_H

#ifdef __cplusplus
extern "C" {
#endif

#define RCMPS_OK 0
#define RCMPS_ERR_MAX_STEPS 1        /* ODE exceeded maxiters attempts */
#define RCMPS_ERR_STEP_UNDERFLOW 2   /* ODE step below 10 ulp */
#define RCMPS_ERR_NONFINITE 3        /* nonfinite error estimate or result */
#define RCMPS_ERR_MISSING_SAMPLES 5  /* trajectory ended before every quadrature node */
#define RCMPS_ERR_WAIT_LIMIT 7       /* a workgroup waited past its time budget (kernel aborted) */
#define RCMPS_ERR_SOLVER 10          /* Krylov or factorisation failure (SolverFailure) */
#define RCMPS_ERR_ARGUMENT 11        /* invalid argument (ArgumentError) */
#define RCMPS_ERR_DEVICE 12          /* HIP runtime error */
#define RCMPS_ERR_UNSUPPORTED 13     /* configuration not built (e.g. D above the kernel limit) */
#define RCMPS_ERR_SCRATCH_LIMIT 14   /* input larger than a no-alloc entry point's fixed scratch */

#define RCMPS_FLAG_NO_BOUNDARY 1     /* research switch of evaluate(...; boundary=false) */
/* Split precision: the D x D products of the trajectory right-hand sides run on the INT8 MFMA
   over exact 8-bit digit slices of the FP64 operands (48-bit operands, exact INT32 sums, FP64
   recombination); integration state, error control, dense output, pullback, boundary and
   metric solves stay FP64. */
#define RCMPS_FLAG_SPLIT 2
/* kinetic_backend=:krylov (src/observables.jl:229-236): both kinetic terms on the infinite line by
   the angular resolvent with matrix-free Arnoldi (the context's rcmps_kinetic_options); only the
   :phi^4: trajectories run. */
#define RCMPS_FLAG_KINETIC_KRYLOV 4
/* fill rcmps_stats.phase_us (the trajectory kernel's per-phase timers); without it the kernel reads
   no clock and phase_us stays zero */
#define RCMPS_FLAG_PHASE_TIMERS 8
/* ignore a state's Z2 structure: every product over every k-tile (the same bits as the default,
   which skips the k-tiles of the parity blocks that are exactly zero when R is odd and K, rho0 even
   under the split of the basis into halves of whole 16 x 16 tiles, D = 32, 64) */
#define RCMPS_FLAG_DENSE 16

typedef struct rcmps_context rcmps_context;
typedef struct rcmps_state rcmps_state;
typedef struct rcmps_excitation rcmps_excitation;

/* parts of rcmps_excitation_action (0 = both) */
#define RCMPS_EXC_KINETIC 1       /* the kinetic terms on the theta grid */
#define RCMPS_EXC_POTENTIAL 2     /* the :phi^4: term by the action ODEs */
#define RCMPS_EXC_DENSE 4         /* with either: ignore a Z2 state's parity blocks in the action ODE
                                     (the same bits as the default, which skips the exactly zero ones
                                     when every tangent has one parity and D = 32, 64) */
#define RCMPS_EXC_GRAPH 8         /* with either: step the action ODE by replayed launch graphs */
#define RCMPS_EXC_PERSISTENT 16   /* with either: integrate it in one persistent launch wherever every
                                     tile's workgroup fits on the device at once (neither flag: the
                                     faster of the two for the size; all three give the same bits) */

/* IntegratorParameters (src/state.jl:4-10) plus the DOP853 initial step. */
typedef struct {
    int n_quad;
    double max_point;
    double atol;
    double rtol;
    long long maxiters;
    double initial_step;
} rcmps_integrator;

/* KineticOptions (src/kinetic.jl:9-16) and KineticKrylovOptions (src/kinetic_krylov.jl:8-13). */
typedef struct {
    int minnodes, maxnodes;        /* angular Gauss nodes, doubled from minnodes up to maxnodes */
    double atol, rtol;             /* convergence of energy and finished gradient between doublings */
    double tolerance;              /* true relative residual of every shifted solve */
    double fallback_tolerance;     /* residual accepted at basis exhaustion or maxdim (>= tolerance) */
    int initialdim, maxdim;        /* Arnoldi basis sizes */
} rcmps_kinetic_options;

typedef struct {
    int status;
    int failed_system;                 /* 0..5 = (kinetic, kinetic commutator, phi^4) x (fwd, adj) */
    long long attempts[6], accepted[6], rejected[6], dense_steps[6], rhs_calls[6];
    double term_values[3];             /* kinetic1, kinetic2, <:phi^4:> */
    double trajectory_ms, pullback_ms, boundary_ms, total_ms;
    int krylov_matvecs;
    double krylov_residual;
    /* per system, workgroup 0 of the trajectory kernel, microseconds spent in: wait for stage
       inputs, load inputs, level-1 products, level-2 products before the neighbour tiles, wait for
       neighbour tiles, load them, remaining products and epilogue, integrator bookkeeping,
       step end (error norm, record, all-gather), whole kernel, publishing the next stage input,
       stage entry.  The excitation action: [0][0..3] its kinetic phases; [4][0..3] workgroup 0 of
       the persistent action ODE launch in the right-hand-side products, the exchanges between
       stages, the step controller and the whole launch */
    double phase_us[6][12];
    /* infinite-line kinetic backend: angular nodes of the accepted level, basis sizes per
       (op R/[Q,R]) x (forward/adjoint), operator applications, largest true relative residual */
    int kinetic_nodes;
    int kinetic_basis[4];
    int kinetic_actions;
    double kinetic_residual;
    /* infinite-line kinetic backend, wall milliseconds: Arnoldi steps, projected shifted solves,
       forming the node solutions and their true residuals, the rest (node sums, gradient finish) */
    double kinetic_ms[4];
} rcmps_stats;

int rcmps_context_create(int device, rcmps_context** out);
void rcmps_context_destroy(rcmps_context* ctx);
const char* rcmps_status_string(int status);
/* largest bond dimension the trajectory kernels are built for */
int rcmps_max_bond_dimension(void);

/* State(R, K): K must be Hermitian (|K - K^+| <= 1e-12 max(1,|K|)); Q = -iK - R^+R/2.  rho0 is
   solved to steady_tol unless `rho` is given, in which case it is checked (checked_density). */
int rcmps_state_create(rcmps_context* ctx, int D, const double* R, const double* K, const double* rho,
                       double steady_tol, rcmps_state** out, rcmps_stats* stats);
void rcmps_state_destroy(rcmps_state* st);
/* copy Q (which = 0) or rho0 (which = 1) into a device D x D buffer */
int rcmps_state_get(rcmps_context* ctx, const rcmps_state* st, int which, double* dst);

/* fg_phi4: energy and Riemannian gradient (metric 2 Re tr(W1^+ W2 (rho0 + metric_reg I))).
   grad, raw_cov, raw_env may be null.  raw_cov/raw_env are the summed covector and
   environment before finish_gradient. */
int rcmps_fg_phi4(rcmps_context* ctx, const rcmps_state* st, double g, const rcmps_integrator* ip,
                  double metric_reg, int flags, double* energy, double* grad, double* raw_cov,
                  double* raw_env, rcmps_stats* stats);

/* fg_phi4 over n independent states (any mix of bond dimensions and couplings g[i]) with one
   integrator: energies[i] (host) and, where grads and grads[i] are non-null, the Riemannian
   gradient into the device buffer grads[i] (D_i x D_i).  States of the same padded size share a
   persistent launch in which their trajectories are gang-scheduled in rounds of resident
   workgroups (the largest size in two launches); each launch's pullback, boundary and metric
   solves run on worker streams while the next launch integrates.  stats may be null or an array
   of n.  Returns the first failure (each stats[i]
   holds its own status); the evaluations that succeeded are complete.  Replaces a loop of
   fg_phi4(s_i, g_i) calls (src/observables.jl:257); the reference has no batched form. */
int rcmps_fg_phi4_batch(rcmps_context* ctx, int n, const rcmps_state* const* states, const double* g,
                        const rcmps_integrator* ip, double metric_reg, int flags, double* energies,
                        double* const* grads, rcmps_stats* stats);

/* Energy and its Euclidean gradient in (R, K), for Q = -iK - R^+R/2 with rho0 re-solved:
   dE = Re sum_ij [conj(grad_R_ij) dR_ij + conj(grad_K_ij) dK_ij] for every dR and every
   Hermitian dK (grad_K is Hermitian), i.e. torch's convention for a real function of complex
   inputs.  No counterpart in the reference, which returns only the Riemannian gradient of
   fg_phi4 (src/observables.jl:161-212); it is the same adjoint pullback with the coefficients
   of dR and dQ kept apart, for callers that differentiate through (R, K) (a
   torch.autograd.Function).  grad_R or grad_K may be null; when both are, only the energy is
   computed and only the three forward trajectories run. */
int rcmps_fg_phi4_rk(rcmps_context* ctx, const rcmps_state* st, double g, const rcmps_integrator* ip, int flags,
                     double* energy, double* grad_R, double* grad_K, rcmps_stats* stats);

/* Set the infinite-line kinetic controls of this context (defaults: 32, 512, 1e-11, 1e-10, 1e-12,
   1e-12, 16, 512).  RCMPS_ERR_ARGUMENT for invalid controls (ArgumentError in the reference). */
int rcmps_set_kinetic_options(rcmps_context* ctx, const rcmps_kinetic_options* opt);

/* Layout of every D x D matrix passed through this context's entry points (states, gradients,
   covectors, environments, boundary and excitation operands; not the padded trajectory samples):
   RCMPS_LAYOUT_ROW_MAJOR (default) or RCMPS_LAYOUT_COLUMN_MAJOR, Julia's own layout and that of
   the portable checkpoint's .bin files (gpu/src/rcmps_gpu/io.py:374,467), so a Julia caller passes
   its arrays as they are instead of permutedims copies.  The device works in padded row-major either way; the copies at the
   boundary transpose.  RCMPS_ERR_ARGUMENT for another value. */
#define RCMPS_LAYOUT_ROW_MAJOR 0
#define RCMPS_LAYOUT_COLUMN_MAJOR 1
int rcmps_set_layout(rcmps_context* ctx, int layout);

/* The no-jump inverse that preconditions the stationary-state, boundary and excitation solves
   (src/transfer.jl:50-68, NoJumpInverse: a Schur-form trsyl solve) runs in the eigenbasis of Q
   (elementwise middle step) unless the eigenvectors are ill-conditioned, as for a defective Q
   (test/runtests.jl:74-87), where it keeps the Schur form and solves the triangular Sylvester
   system by an anti-diagonal sweep (D <= 64).  States created on `ctx` take the Schur form when
   the condition estimate |W|_F |W^-1|_F of Q's triangular eigenvectors exceeds `kappa` (default
   1e6; 0 always, INFINITY never).  The Krylov solutions are exact either way.
   RCMPS_ERR_ARGUMENT for a negative or NaN kappa. */
int rcmps_set_nojump_threshold(rcmps_context* ctx, double kappa);
/* The form a state's no-jump inverse took: *schur = 1 (Schur basis) or 0 (eigenbasis), *kappa =
   the eigenvector condition estimate (INFINITY when W is singular). */
int rcmps_state_nojump_form(const rcmps_state* st, int* schur, double* kappa);

/* The eigendecomposition L = W diag(mu) W^-1 of the state's transfer superoperator
   L(X) = Q X + X Q^+ + R X R^+ on row-major vec(X) (index i D + j), n = D^2: W and W^-1 column-major
   n x n, mu n eigenvalues, device pointers (copied).  With it the excitation kinetic resolvents
   (1 - z (L - ip))^-1 and (1 - z (L* + ip))^-1 of every momentum become products with W and W^-1
   (the reference's ExcitationContext factors the same transfer densely,
   src/excitation_environments.jl:18-40) instead of Krylov bases; n = 0 removes it.  Computed on the
   host by the caller (rcmps_gpu does it for D <= 32).  RCMPS_ERR_ARGUMENT for another n or
   nonfinite entries. */
int rcmps_state_set_transfer_eigen(rcmps_context* ctx, rcmps_state* st, int n, const double* W, const double* Winv,
                                   const double* mu);

/* Tile variant of this context's batched complex products: -1 (default) lets the autotuner pick
   per (padded size, batch class), 0..3 force one (0: one 16x16 tile per four-wave workgroup with
   the k-steps split, 1: four tiles per workgroup, one per wave, 2: one tile per one-wave
   workgroup, 3: a 32x32 block per workgroup from operand panels staged in LDS).  Every variant
   returns the same bits.  No counterpart in the reference, whose
   products are library GEMMs; the tuner is the batching path's tile-shape choice.
   RCMPS_ERR_ARGUMENT for another value. */
int rcmps_gemm_variant(rcmps_context* ctx, int variant);

/* The autotuner's table: for up to `cap` entries, the padded size, batch class c (batches of
   2^(c-1) < n <= 2^c), the chosen variant and the measured microseconds per launch of each of
   the four variants (us[4 i + v]); *count receives the number of entries.  No counterpart in
   the reference. */
int rcmps_gemm_tuning(rcmps_context* ctx, int cap, int* padded_size, int* batch_class, int* variant, double* us,
                      int* count);

/* kinetic1 (which = 1), kinetic2 (which = 2) or both (which = 3) on the infinite line by the Krylov
   backend: kinetic_resolvent (src/kinetic.jl:77-124) over kinetic_components
   (src/kinetic_krylov.jl:156-189).  raw_cov, raw_env (device, may be null) are the covector and
   environment before finish_gradient; grad (device, may be null) the finished Riemannian gradient
   (metric_reg as in rcmps_fg_phi4). */
int rcmps_kinetic(rcmps_context* ctx, const rcmps_state* st, int which, double metric_reg, double* value,
                  double* raw_cov, double* raw_env, double* grad, rcmps_stats* stats);

/* Solve (-L* + |I><rho0|) H = a - tr(rho0 a) I, tr(rho0 H) = 0. */
int rcmps_boundary_environment(rcmps_context* ctx, const rcmps_state* st, const double* a, double tol,
                               double* h, rcmps_stats* stats);

/* rcmps_boundary_environment without any allocation: the same solve with the same kernels and the
   same result, its device workspace (GMRES basis, temporaries, device scalars) carved from a
   library-owned static device array of 4 MiB per GPU (held until the process ends) and its
   Hessenberg read-back through host stack memory.  Fits every D <= 64 (D = 64 takes 3.2 MiB);
   a state too large for the array returns RCMPS_ERR_SCRATCH_LIMIT.  All calls on one GPU share its
   array, so calls that overlap on one GPU race (serialise them, e.g. one context per GPU).  No
   counterpart in the reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The context must
   have been created (rcmps_context_create sets the array up on its device). */
int rcmps_boundary_environment_noalloc(rcmps_context* ctx, const rcmps_state* st, const double* a, double tol,
                                       double* h, rcmps_stats* stats);

/* After rcmps_fg_phi4 on this context: copy the samples of one trajectory system
   (2 n_quad + 1 points x live channels x D x D, row-major) into a device buffer of
   `capacity` doubles.  Returns RCMPS_ERR_ARGUMENT if it is too small. */
int rcmps_trajectory_samples(rcmps_context* ctx, int system, double* dst, long long capacity);

/* ExcitationContext of a state at momentum p: the tangent-space excitation action works on it.  The
   state must outlive the context.  No D^2 x D^2 transfer matrix is formed: the Poisson and resolvent
   solves are matrix-free (GMRES and shifted Arnoldi), so the reference's dense mixing check is
   replaced by the residual checks of those solves (a non-mixing state fails them with
   RCMPS_ERR_SOLVER).  environment_tol is the reference's `tol` (1e-10). */
int rcmps_excitation_create(rcmps_context* ctx, const rcmps_state* st, double momentum, double environment_tol,
                            rcmps_excitation** out);
void rcmps_excitation_destroy(rcmps_excitation* ex);

/* adjoint = 0: x solves (-L + i shift p + |rho><I|) x = b - tr(b) rho, projected to tr(x) = 0
   (excitation_poisson); adjoint = 1: (-L* - i shift p + |I><rho|) x = b - tr(rho b) I, projected to
   tr(rho x) = 0 (action_dual_poisson).  shift in {-1, 0, 1}; b, x device D x D. */
int rcmps_excitation_poisson(rcmps_context* ctx, rcmps_excitation* ex, const double* b, int adjoint, int shift,
                             double* x);

/* F_i = H(W_i) for n tangents W_i (device D x D; F_i device D x D) with <U, F(W)> the connected matrix
   element <Phi_p(U)|H - E0|Phi_p(W)> per 2 pi delta(0).  The model is phi^4: kinetic_coef[0] (op R) and
   kinetic_coef[1] (op [Q, R]) weight the kinetic terms (null: 1, 1) on the fixed theta grid of
   kinetic_nodes Gauss nodes, and g weights :phi^4: integrated with `ip`.  `parts` selects
   RCMPS_EXC_KINETIC and/or RCMPS_EXC_POTENTIAL (0: both; the kinetic preconditioner uses the kinetic
   part alone).  All tangents of a call share their Krylov iterations and ODE steps. */
int rcmps_excitation_action(rcmps_context* ctx, rcmps_excitation* ex, int n, const double* const* W, double* const* F,
                            double g, const double* kinetic_coef, const rcmps_integrator* ip, int kinetic_nodes,
                            int parts, rcmps_stats* stats);

/* rcmps_excitation_action with one context per tangent: contexts of different momenta on one state, so
   that the Davidson searches of several momenta (iterative_dispersion, src/excitation_eigsolve.jl:453-476)
   apply their actions in one call.  All exs[i] must belong to the same state (RCMPS_ERR_ARGUMENT
   otherwise).  No counterpart in the reference, which applies one W at one momentum per call. */
int rcmps_excitation_action_multi(rcmps_context* ctx, int n, rcmps_excitation* const* exs, const double* const* W,
                                  double* const* F, double g, const double* kinetic_coef, const rcmps_integrator* ip,
                                  int kinetic_nodes, int parts, rcmps_stats* stats);

/* Connected matrix elements <Phi_p(U_i)|H - E0|Phi_p(W_i)> per 2 pi delta(0) of n tangent pairs (each with
   its context; one state): the pair environments, the four-block :phi^4: transfer ODE and the kinetic
   resolvent sum on kinetic_nodes theta nodes (excitation_element, src/excitations.jl:98-113, with
   ExcitationPair src/excitation_environments.jl:105-127).  out (host): 2n doubles, the complex values.
   The route the reference's validation calls "bilinear" (src/excitation_eigsolve.jl:273-279). */
int rcmps_excitation_element(rcmps_context* ctx, int n, rcmps_excitation* const* exs, const double* const* U,
                             const double* const* W, double g, const double* kinetic_coef, const rcmps_integrator* ip,
                             int kinetic_nodes, int parts, double* out, rcmps_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
