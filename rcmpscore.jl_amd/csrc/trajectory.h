// Shared host/device description of the persistent DOP853 trajectory launch.
#pragma once
#include <stdint.h>

#define RC_NSYS 6           // (kinetic, kinetic commutator, phi^4) x (forward, adjoint)
#define RC_MAX_LIVE 4
#define RC_NSTAGE_BUF 16    // one buffer per stage slot of a step
#define RC_REC 16           // doubles per workgroup record: e5, e3, 12 trace partials, spare

// status codes (shared with the C-ABI): 1-5 as the DOP853 controller of the reference GPU
// package (device_control_backend.py:23-27), 7 = a workgroup waited past its time budget.
#define RC_ST_OK 0
#define RC_ST_MAX_STEPS 1
#define RC_ST_STEP_UNDERFLOW 2
#define RC_ST_NONFINITE 3
#define RC_ST_MISSING_SAMPLES 5
#define RC_ST_WAIT_LIMIT 7

struct TrajTask {
    int sys, ch, b, c;
    int adj;
    int edge_src;       // live channel (same system) whose level-1 tile feeds the insertion, -1 none
    int edge_from_E;    // 1: the insertion reads the E buffer (X * [Q,R]) of edge_src
    int has_E;          // load the C = [Q, R] column block of E = X C (the E tasks of the kinetic commutator)
    int trace_kind;     // 0 none; 1: Re sum X conj(O) over the tile; 2: Re diag of the level-1 tile
    int e_only;         // 1: only form and publish channel ch's E tile each stage (no stage input,
                        // no samples, an empty error record)
    // k-tiles that contribute to each tile product, four bits per product: level 1 [0, 4), the two
    // level-2a products [4, 8) and [8, 12), level 2b [12, 16), the E product [16, 20).  All ones
    // for a general state; for a Z2-symmetric state only the k-tiles of the one parity block that
    // can be nonzero (none when the product is structurally zero): the skipped k-steps multiply
    // exact zeros, so the results are the same bits.
    int kmask;
    double edge_coef;
    double trace_coef;
    const double* fixed;     // Dp x Dp complex (row-major, interleaved) or null
    const double* trace_op;  // O for trace kind 1
};

struct TrajSys {
    int nlive, has_scalar, adj, first_task, ntasks, D;
    double n_elems;          // nlive * D^2 + has_scalar: denominator of the RMS norm
    const double *Q, *Qh, *R, *Rh, *C;   // the state's matrices, Dp x Dp
    double* Y;               // [16][nlive][Dp*Dp] complex
    double* P;               // [16][nlive][Dp*Dp] complex
    double* E;               // [16][Dp*Dp] complex or null
    double* rec;             // [2][ntasks][RC_REC]
    uint32_t* flagY;         // [nlive][NB*NB]
    uint32_t* flagP;         // [nlive][NB*NB]
    uint32_t* flagE;         // [NB*NB]
    uint32_t* flagR;         // [ntasks]
    double* samples;         // [npoints][nlive][Dp*Dp] complex
    double* out;             // [32]: status/counters, then phase timers of workgroup 0
};

// One launch integrates the systems of one or more states of the same padded size Dp.  Tasks
// are gang-scheduled in rounds: workgroup w runs task schedule[r * nblocks + w] (or idles, -1)
// in round r, and all tasks of a system share a round, so every wait is on a running partner.
struct TrajParams {
    int Dp, npoints, nrounds, timers;   // timers: fill the per-phase timers of workgroup 0
    double a, atol, rtol, h0;
    long long max_steps, wait_ticks;   // wait_ticks: time budget of one wait (constant-rate clock ticks)
    const double* points;
    const double* cheb;
    const TrajTask* tasks;
    const TrajSys* sys;
    const int* schedule;     // [nrounds][nblocks]
    uint32_t* abort_flag;
};

// out[] layout per system
enum { RC_OUT_STATUS = 0, RC_OUT_ATTEMPTS, RC_OUT_ACCEPTED, RC_OUT_REJECTED, RC_OUT_DENSE,
       RC_OUT_RHS, RC_OUT_SCALAR, RC_OUT_T, RC_OUT_NEXTPOINT, RC_OUT_TIMERS = 16 };
#define RC_OUT_STRIDE 32
#define RC_NTIMERS 12
