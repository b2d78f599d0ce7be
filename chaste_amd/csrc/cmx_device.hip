// MI355X (gfx950) monodomain time-step: device kernels and the C ABI.
//
// One PDE step (AbstractDynamicLinearPdeSolver::Solve body,
// pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:372-469) is the launch
// sequence
//   cell_step    SolveCellSystems(V, t, t+dt) with V clamped, fused with the
//                ICI forcing F_i = -Am Iion_i - Istim_i(t+dt)
//                (heart/src/tissue/AbstractCardiacTissue.cpp:464-734,
//                 heart/src/solver/electrics/monodomain/MonodomainSolver.cpp:92-101)
//   residual     r0 = b - A V^n with b = M (Am Cm V^n/dt + F) and
//                A = (Am Cm/dt) M + K, evaluated as r0 = M F - K V^n (the
//                Am Cm V/dt terms cancel exactly), plus ||D^-1 b|| for the test
//                (MonodomainSolver.cpp:107 MatMult, LinearSystem::Solve)
//   cg_start     w0 = A D^-1 r0 and the first dot products
//   cg_iter x B  Jacobi-preconditioned CG on A d = r0 from d = 0, which is the
//                reference's warm start from V^n (KSPSetInitialGuessNonzero,
//                linalg/src/LinearSystem.cpp:833-837).  Pipelined form
//                (Ghysels & Vanroose): one kernel and one reduction per
//                iteration; stops when ||D^-1 r|| <= max(rtol ||D^-1 b||, atol)
//                (PETSc's default preconditioned-norm test); iterations past
//                convergence are no-ops
// and the next step's cell kernel (or a batch's finish kernel) settles it:
// the last reduction decides the solve, V^{n+1} = V^n + d in FP64, and the
// device step clock advances.  CMX_MIXED keeps cell states, F and the CG
// vectors in FP32 and adds an FP64 refinement (true residual of V^n + d, a
// second CG solve when that fails the test); V, residuals, dot products and
// the voltage update are FP64 throughout.  The true residual is computed by
// the first solve's last budgeted iteration (cg_last) when the refinement
// solve has no budget, else by a kernel of its own (refine_residual).
//
// Steps are enqueued in batches without host synchronisation.  If a step's
// final solve is still running when its iteration budget B is used up, the
// settling kernel poisons the clock, every later kernel returns at once, and
// the host resumes that solve where it stopped (all its buffers are intact),
// then re-enqueues the rest of the batch.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <dlfcn.h>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "cmx_models.h"

namespace cmx {

// ---------------------------------------------------------------------------
// Linear algebra
// ---------------------------------------------------------------------------
struct MatVals {
    const double* __restrict__ m64;   // mass
    const double* __restrict__ k64;   // stiffness
    const double* __restrict__ a64;   // (Am Cm/dt) M + K
    const float* __restrict__ a32;
    const double* __restrict__ dinv64;
    const float* __restrict__ dinv32;
};

template <class RV> __device__ __forceinline__ const RV* a_vals(const MatVals& m);
template <> __device__ __forceinline__ const double* a_vals<double>(const MatVals& m) { return m.a64; }
template <> __device__ __forceinline__ const float* a_vals<float>(const MatVals& m) { return m.a32; }
template <class RV> __device__ __forceinline__ const RV* dinv_vals(const MatVals& m);
template <> __device__ __forceinline__ const double* dinv_vals<double>(const MatVals& m) { return m.dinv64; }
template <> __device__ __forceinline__ const float* dinv_vals<float>(const MatVals& m) { return m.dinv32; }


// CG work vectors (RV precision).  The stencil gathers only p (which holds
// u0 = D^-1 r0 before the first iteration) and m; every other vector is read
// and written by its own row only.
template <class RV>
struct CgVecs {
    RV* r;       // residual
    RV* w;       // A u with u = D^-1 r
    RV* z;       // A q with q = D^-1 s
    RV* s;       // A p
    RV* p;       // search direction
    RV* m_in;    // D^-1 w of the previous iteration (gathered)
    RV* m_out;   // D^-1 w after this iteration
    RV* x;       // correction being solved for
    RV* u_out = nullptr;   // D^-1 r after this iteration (for the folded refinement that follows), or null
};

// Every solver kernel below issues all of its global loads (tables, the
// previous reduction, the rows' own and gathered values) before its first
// barrier, so one memory round trip covers them; the iteration kernel is the
// exception (it decides from the reduction first, so that launches after
// convergence cost no row traffic).

template <class RV>
struct ResidualArgs {
    int n;
    double c;                          // Am Cm / dt
    StepClock* clock;
    int32_t* its_log;                  // step statistics (first solve's residual kernel)
    int log_len;
    int refined;
    const double* __restrict__ Z;      // V^n, or V^n + x0 with an initial guess
    const RV* __restrict__ G;          // F, or F - c x0 with an initial guess
    const RV* __restrict__ svi;        // state-variable interpolation correction of b, or null
    double* __restrict__ r0_64;        // FP64 residual, kept for the refinement sweep
    const RV* __restrict__ x1;         // refinement sweep: the first solve's correction
    const double* __restrict__ pin0;   // refinement sweep: the first solve's last reduction
    int gin0;                          //   its slot count
    SolveCtl* ctl0;                    //   and the first solve (its convergence is recorded)
    CgVecs<RV> v;
    double* __restrict__ partials;     // [G][3]: ||D^-1 b||^2, 0, ||D^-1 r0||^2
    SolveCtl* ctl;
};

// r0 = b - A Z for the solve A d = r0 (V^{n+1} = Z + d) with
// b = M (c V^n + F) (+ the state-variable interpolation correction S,
// MonodomainSolver.cpp:122-127) and A = c M + K (c = Am Cm / dt).  Z = V^n is the
// reference's warm start (KSPSetInitialGuessNonzero from V^n); with an
// initial guess Z = V^n + x0 and G = F - c x0, otherwise G = F, and in both
// cases r0 = M G - K Z and b = M (G + c Z).  u0 = D^-1 r0 goes to p for the
// solve that follows, d = 0, and r0 is also kept in FP64 for the refinement
// sweep.  The stiffness term uses differences (Z_j - Z_i): K's rows sum to
// zero, and the differences keep the diffusion flux free of cancellation
// against the ~85 mV resting level.
template <class Rows, class RV>
struct ResidualRow {
    typename Rows::Row row;
    int i;
    double zi, si;
    Gathered<Rows, double> Z;
    Gathered<Rows, RV> G;
    __device__ __forceinline__ void load(const Rows& R, int i_, const ResidualArgs<RV>& a)
    {
        i = i_;
        row = R.row(i);
        zi = a.Z[i];
        si = a.svi ? (double)a.svi[i] : 0.0;
        Z.load(R, row, a.Z);
        G.load(R, row, a.G);
    }
    __device__ __forceinline__ void apply(const Rows& R, const double* __restrict__ Mv, const double* __restrict__ Kv,
                                          const double* __restrict__ Dv, const RV* __restrict__ D32,
                                          const ResidualArgs<RV>& a, double (&part)[3]) const
    {
        double mg = 0, mz = 0, kz = 0;
#pragma unroll
        for (int k = 0; k < R.width(); ++k) {
            const int s = R.vidx(row, k);
            const double zj = Z.at(R, row, a.Z, k);
            mg += Mv[s] * (double)G.at(R, row, a.G, k);
            mz += Mv[s] * zj;
            kz += Kv[s] * (zj - zi);
        }
        const int d = R.didx(row);
        const double bi = Dv[d] * (a.c * mz + mg + si);
        const double r64 = mg - kz + si;
        if (a.r0_64) a.r0_64[i] = r64;
        const RV ri = (RV)r64;
        const RV ui = D32[d] * ri;
        a.v.r[i] = ri;
        a.v.p[i] = ui;
        a.v.x[i] = RV(0);
        part[0] += bi * bi;
        part[2] += (double)ui * (double)ui;
    }
};

template <class Rows, class RV>
__global__ __launch_bounds__(kCgBlock) void residual_kernel(Rows R, MatVals mv, ResidualArgs<RV> a)
{
    __shared__ double sM[kMaxPatterns * kMaxStencil], sK[kMaxPatterns * kMaxStencil], sD[kMaxPatterns];
    const TableLoad<Rows, double> tm(mv.m64, R.value_count()), tk(mv.k64, R.value_count()),
        td(mv.dinv64, R.diag_count());
    const int poison = a.clock->poison;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(a.n, blk, gridDim.x);
    const int i0 = rr.lo + threadIdx.x;
    ResidualRow<Rows, RV> q;
    if (i0 < rr.hi) q.load(R, i0, a);
    const double* __restrict__ Mv = tm.store(sM, mv.m64);
    const double* __restrict__ Kv = tk.store(sK, mv.k64);
    const double* __restrict__ Dv = td.store(sD, mv.dinv64);
    if constexpr (Rows::kStaged) __syncthreads();
    if (poison) return;
    const RV* __restrict__ D32 = dinv_vals<RV>(mv);
    double part[3] = {0.0, 0.0, 0.0};
    if (i0 < rr.hi) q.apply(R, Mv, Kv, Dv, D32, a, part);
    for (int i = i0 + kCgBlock; i < rr.hi; i += kCgBlock) {
        ResidualRow<Rows, RV> q2;
        q2.load(R, i, a);
        q2.apply(R, Mv, Kv, Dv, D32, a, part);
    }
    write_partials<kCgBlock, 3>(part, a.partials, blk);
    if (blk == 0 && threadIdx.x == 0) {
        // the previous step was settled by this step's cell kernel: count it
        StepClock* ck = a.clock;
        if (ck->pending) {
            log_step(a.its_log, a.log_len, ck->k, a.ctl, a.refined != 0);
            ck->k += 1;
        }
        ck->pending = 1;
        a.ctl->done = 0;
        a.ctl->it = 0;
    }
}

// Refinement sweep: the true FP64 residual of V^n + x1 against the step's b,
// r = r0 - A x1 (A in FP64), then u = D^-1 r to p and x2 = 0 for the second
// solve.  Partials: (0, 0, ||D^-1 r||^2).
template <class Rows, class RV>
struct RefineRow {
    typename Rows::Row row;
    int i;
    double r0;
    Gathered<Rows, RV> X;
    __device__ __forceinline__ void load(const Rows& R, int i_, const ResidualArgs<RV>& a)
    {
        i = i_;
        row = R.row(i);
        r0 = a.r0_64[i];
        X.load(R, row, a.x1);
    }
    __device__ __forceinline__ void apply(const Rows& R, const double* __restrict__ Av, const RV* __restrict__ D32,
                                          const ResidualArgs<RV>& a, double (&part)[3]) const
    {
        double ax = 0;
#pragma unroll
        for (int k = 0; k < R.width(); ++k) ax += Av[R.vidx(row, k)] * (double)X.at(R, row, a.x1, k);
        const RV ri = (RV)(r0 - ax);
        const RV ui = D32[R.didx(row)] * ri;
        a.v.r[i] = ri;
        a.v.p[i] = ui;
        a.v.x[i] = RV(0);
        part[2] += (double)ui * (double)ui;
    }
};

template <class Rows, class RV>
__global__ __launch_bounds__(kCgBlock) void refine_residual_kernel(Rows R, MatVals mv, ResidualArgs<RV> a)
{
    __shared__ double sA[kMaxPatterns * kMaxStencil];
    const TableLoad<Rows, double> ta(mv.a64, R.value_count());
    const int poison = a.clock->poison;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(a.n, blk, gridDim.x);
    const int i0 = rr.lo + threadIdx.x, i1 = i0 + kCgBlock;
    RefineRow<Rows, RV> q0, q1;
    if (i0 < rr.hi) q0.load(R, i0, a);
    if (i1 < rr.hi) q1.load(R, i1, a);
    // block 0 also decides whether the first solve met its own test, for the
    // iteration-budget statistics (the iteration that would have detected it
    // is not launched when the budget is exact)
    PartialLoad<3, kCgBlock> pl0(a.pin0, blk == 0 ? a.gin0 : 0);
    const int done0 = (blk == 0 && threadIdx.x == 0) ? a.ctl0->done : 1;
    const double tol0 = (blk == 0 && threadIdx.x == 0) ? a.ctl0->tol2 : 0.0;
    const double* __restrict__ Av = ta.store(sA, mv.a64);
    if constexpr (Rows::kStaged) __syncthreads();
    if (poison) return;
    const RV* __restrict__ D32 = dinv_vals<RV>(mv);
    double part[3] = {0.0, 0.0, 0.0};
    if (i0 < rr.hi) q0.apply(R, Av, D32, a, part);
    if (i1 < rr.hi) q1.apply(R, Av, D32, a, part);
    for (int i = i1 + kCgBlock; i < rr.hi; i += kCgBlock) {
        RefineRow<Rows, RV> q;
        q.load(R, i, a);
        q.apply(R, Av, D32, a, part);
    }
    write_partials<kCgBlock, 3>(part, a.partials, blk);
    if (blk == 0) {
        double g0[3];
        pl0.total(g0);
        if (threadIdx.x == 0) {
            (void)solve_converged(a.ctl0, done0, tol0, g0[2], true);
            a.ctl->done = 0;
            a.ctl->it = 0;
        }
    }
}

template <class RV>
struct CgArgs {
    int n;
    int first;                         // first solve of the step: set the tolerance
    double rtol2, atol2;
    const StepClock* clock;
    SolveCtl* ctl;                     // this solve
    SolveCtl* ctl_all;                 // both solves of the step (tolerance is shared)
    double2* hist;                     // [max_its + 1]: (gamma_i, alpha_i)
    const double* __restrict__ pin;    // partials read
    double* __restrict__ pout;         // partials written
    CgVecs<RV> v;
};

// Start of a solve: w0 = A u0, m0 = D^-1 w0 and the dot products
// (r0, u0), (w0, u0), (u0, u0) for iteration 0.  The residual kernel left
// ||D^-1 b||^2 and ||D^-1 r0||^2: the first solve of a step turns the former
// into the stopping threshold, and a residual already below it (the usual
// case of a refinement sweep) ends the solve here with no iteration.
template <class Rows, class RV>
struct StartRow {
    typename Rows::Row row;
    int i;
    RV r, u;
    Gathered<Rows, RV> U;
    __device__ __forceinline__ void load(const Rows& R, int i_, const CgVecs<RV>& v)
    {
        i = i_;
        row = R.row(i);
        r = v.r[i];
        u = v.p[i];
        U.load(R, row, v.p);
    }
    __device__ __forceinline__ void apply(const Rows& R, const RV* __restrict__ A, const RV* __restrict__ D,
                                          const CgVecs<RV>& v, double (&part)[3]) const
    {
        RV w = 0;
#pragma unroll
        for (int k = 0; k < R.width(); ++k) w += A[R.vidx(row, k)] * U.at(R, row, v.p, k);
        v.w[i] = w;
        v.m_out[i] = D[R.didx(row)] * w;
        part[0] += (double)r * (double)u;
        part[1] += (double)w * (double)u;
        part[2] += (double)u * (double)u;
    }
};

template <class Rows, class RV, int ROWS = (sizeof(RV) == 4 ? 2 : 1)>
__global__ __launch_bounds__(kCgBlock) void cg_start_kernel(Rows R, MatVals mv, CgArgs<RV> a)
{
    __shared__ RV sA[kMaxPatterns * kMaxStencil], sD[kMaxPatterns];
    const TableLoad<Rows, RV> ta(a_vals<RV>(mv), R.value_count()), td(dinv_vals<RV>(mv), R.diag_count());
    PartialLoad<3> pl(a.pin, gridDim.x);
    const int poison = a.clock->poison;
    const double tol2_set = a.ctl->tol2;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(a.n, blk, gridDim.x);
    // rows per thread held in registers before the reduction (ROWS, a launch
    // knob: 2 by default in FP32, 1 in FP64 for the register budget); the
    // block's further rows follow in a loop
    constexpr bool kTwo = ROWS == 2;
    const int i0 = rr.lo + threadIdx.x, i1 = i0 + kCgBlock;
    StartRow<Rows, RV> q0, q1;
    if (i0 < rr.hi) q0.load(R, i0, a.v);
    if (kTwo && i1 < rr.hi) q1.load(R, i1, a.v);
    const RV* __restrict__ A = ta.store(sA, a_vals<RV>(mv));
    const RV* __restrict__ D = td.store(sD, dinv_vals<RV>(mv));
    double g0[3];
    pl.total(g0);   // its barriers also cover the table stores
    if (poison) return;
    const double tol2 = a.first ? fmax(a.rtol2 * g0[0], a.atol2) : tol2_set;
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (lead && a.first) {
        a.ctl_all[0].bb = g0[0];
        a.ctl_all[0].tol2 = tol2;
        a.ctl_all[1].bb = g0[0];
        a.ctl_all[1].tol2 = tol2;
    }
    if (g0[2] <= tol2) {
        if (lead) { a.ctl->done = 1; a.ctl->it = 0; a.ctl->zz = g0[2]; }
        return;
    }
    double part[3] = {0.0, 0.0, 0.0};
    if (i0 < rr.hi) q0.apply(R, A, D, a.v, part);
    if (kTwo && i1 < rr.hi) q1.apply(R, A, D, a.v, part);
    for (int i = (kTwo ? i1 : i0) + kCgBlock; i < rr.hi; i += kCgBlock) {
        StartRow<Rows, RV> q;
        q.load(R, i, a.v);
        q.apply(R, A, D, a.v, part);
    }
    write_partials<kCgBlock, 3>(part, a.pout, blk);
}

// Iteration i of the pipelined preconditioned CG of Ghysels & Vanroose
// (Parallel Computing 40:224, 2014, Alg. 4) with the Jacobi preconditioner
// applied pointwise (u = D^-1 r, q = D^-1 s, m = D^-1 w need no vectors of
// their own except m, which the stencil gathers).  One kernel per iteration:
// the dot products gamma_i = (r_i, u_i), delta_i = (w_i, u_i) and
// zeta_i = ||u_i||^2 come reduced from the previous kernel, and the one
// matrix product n_i = A m_i does not depend on them:
//   beta_i = gamma_i / gamma_{i-1},
//   alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})   (i > 0)
//   z = n + beta z;  s = w + beta s;  p = u + beta p;
//   x += alpha p;  r -= alpha s;  w -= alpha z;  m = D^-1 w
// and the new dot products go to the next kernel.  The iteration stops (x
// final) when zeta_i <= tol2, PETSc's preconditioned-norm test; the kernel
// that consumes the solve applies the same test to the last reduction.
template <class Rows, class RV>
struct IterRow {
    typename Rows::Row row;
    int i;
    RV r, w, x, z, s, p;
    Gathered<Rows, RV> M;
    __device__ __forceinline__ void load(const Rows& R, int i_, const CgVecs<RV>& v, bool first_it)
    {
        i = i_;
        row = R.row(i);
        M.load(R, row, v.m_in);
        r = v.r[i];
        w = v.w[i];
        x = v.x[i];
        if (!first_it) {
            z = v.z[i];
            s = v.s[i];
            p = v.p[i];
        }
    }
    __device__ __forceinline__ void apply(const Rows& R, const RV* __restrict__ A, const RV* __restrict__ D,
                                          const CgVecs<RV>& v, bool first_it, RV al, RV be, double (&part)[3]) const
    {
        RV n = 0;
#pragma unroll
        for (int k = 0; k < R.width(); ++k) n += A[R.vidx(row, k)] * M.at(R, row, v.m_in, k);
        const RV d = D[R.didx(row)];
        RV zn, sn, pn;
        if (first_it) {
            zn = n;
            sn = w;
            pn = d * r;
        } else {
            zn = n + be * z;
            sn = w + be * s;
            pn = d * r + be * p;
        }
        const RV rn = r - al * sn;
        const RV wn = w - al * zn;
        const RV un = d * rn;
        v.z[i] = zn;
        v.s[i] = sn;
        v.p[i] = pn;
        v.x[i] = x + al * pn;
        v.r[i] = rn;
        v.w[i] = wn;
        v.m_out[i] = d * wn;
        if (v.u_out) v.u_out[i] = un;
        part[0] += (double)rn * (double)un;
        part[1] += (double)wn * (double)un;
        part[2] += (double)un * (double)un;
    }
};

template <class Rows, class RV, int ROWS = (sizeof(RV) == 4 ? 2 : 1)>
__global__ __launch_bounds__(kCgBlock, Rows::kIterWaves) void cg_iter_kernel(Rows R, MatVals mv, CgArgs<RV> a, int it)
{
    __shared__ RV sA[kMaxPatterns * kMaxStencil], sD[kMaxPatterns];
    const TableLoad<Rows, RV> ta(a_vals<RV>(mv), R.value_count()), td(dinv_vals<RV>(mv), R.diag_count());
    PartialLoad<3> pl(a.pin, gridDim.x);
    SolveCtl* ctl = a.ctl;
    const int poison = a.clock->poison;
    const int done = ctl->done;
    const double tol2 = ctl->tol2;
    const int max_its = ctl->max_its;
    // a launch after the solve has ended returns before the reduction
    // (uniform: every thread read the same flags)
    if (poison || done) return;
    const RV* __restrict__ A = ta.store(sA, a_vals<RV>(mv));
    const RV* __restrict__ D = td.store(sD, dinv_vals<RV>(mv));
    double g[3];
    pl.total(g);    // its barriers also cover the table stores
    const double gamma = g[0], delta = g[1], zeta = g[2];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (!(zeta == zeta) || !(gamma == gamma) || !(delta == delta) || fabs(zeta) > DBL_MAX) {
        if (lead) { ctl->done = 3; ctl->it = it; ctl->zz = zeta; }
        return;
    }
    if (zeta <= tol2) {
        if (lead) { ctl->done = 1; ctl->it = it; ctl->zz = zeta; }
        return;
    }
    if (it >= max_its) {
        if (lead) { ctl->done = 2; ctl->it = it; ctl->zz = zeta; }
        return;
    }
    const RowRange rr(a.n, blk, gridDim.x);
    const bool first_it = it == 0;
    constexpr bool kTwo = ROWS == 2;   // rows per thread held in registers (launch knob)
    const int i0 = rr.lo + threadIdx.x, i1 = i0 + kCgBlock;
    IterRow<Rows, RV> q0, q1;
    if (i0 < rr.hi) q0.load(R, i0, a.v, first_it);
    if (kTwo && i1 < rr.hi) q1.load(R, i1, a.v, first_it);
    double alpha, beta;
    if (first_it) {
        beta = 0.0;
        alpha = gamma / delta;
    } else {
        const double2 h = a.hist[it - 1];
        beta = gamma / h.x;
        alpha = gamma / (delta - beta * gamma / h.y);
    }
    if (lead) {
        a.hist[it] = make_double2(gamma, alpha);
        ctl->it = it + 1;
        ctl->zz = zeta;
    }
    const RV al = (RV)alpha, be = (RV)beta;
    double part[3] = {0.0, 0.0, 0.0};
    if (i0 < rr.hi) q0.apply(R, A, D, a.v, first_it, al, be, part);
    if (kTwo && i1 < rr.hi) q1.apply(R, A, D, a.v, first_it, al, be, part);
    for (int i = (kTwo ? i1 : i0) + kCgBlock; i < rr.hi; i += kCgBlock) {
        IterRow<Rows, RV> q;
        q.load(R, i, a.v, first_it);
        q.apply(R, A, D, a.v, first_it, al, be, part);
    }
    write_partials<kCgBlock, 3>(part, a.pout, blk);
}

// The last budgeted iteration of a refined step's first solve with the FP64
// refinement residual folded in (FP32 vectors).  It makes the iteration's
// x update -- or none, if the solve has already ended -- and in the same
// pass the true residual of the result, r = r0 - A x (A and r0 in FP64),
// which is what decides the step, as the separate refinement residual kernel
// does.  A row needs its neighbours' new x: they are recomputed from values
// this kernel does not write (x, u = D^-1 r from the previous iteration, p),
//   x_new = x + alpha (u + beta p)      (x_new = alpha p at iteration 0),
// with the same rounded operations as the row's own update, so the residual
// is that of exactly the x that is stored.  The result goes to xf (the first
// solve's answer); r, D^-1 r and x = 0 go to the refinement solve's vectors,
// for the host to resume with if the residual fails the test.  Partials: (0,
// ||D^-1 r_rec||^2 of the recursive residual after the update (statistics),
// ||D^-1 r||^2 of the true one).
template <class RV>
struct LastArgs {
    CgArgs<RV> cg;                     // this iteration (cg.pin: the reduction it reads)
    const double* __restrict__ r0_64;  // the step's FP64 residual r0 = b - A Z
    const RV* __restrict__ u_in;       // D^-1 r after the previous iteration (it > 0)
    RV* __restrict__ xf;               // the first solve's final x
    SolveCtl* ctl1;                    // refinement solve (reset here)
    CgVecs<RV> v1;                     // refinement solve: r, p = D^-1 r, x = 0
};

__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }

// x + alpha (u + beta p) with its two roundings fixed (no contraction choice
// left to the compiler), so every thread that forms it gets the same bits.
template <class RV>
__device__ __forceinline__ RV last_x(RV x, RV u, RV p, RV al, RV be)
{
    return fma_rn(al, fma_rn(be, p, u), x);
}

template <class Rows, class RV>
struct LastRow {
    typename Rows::Row row;
    int i;
    double r0;
    RV r, w, s, x, u, p;
    Gathered<Rows, RV> X, U, P;
    // upd: the update's inputs too (the first iteration's are loaded
    // before the decision, so always)
    __device__ __forceinline__ void load(const Rows& R, int i_, const LastArgs<RV>& a, bool first_it, bool upd)
    {
        i = i_;
        row = R.row(i);
        r0 = a.r0_64[i];
        const CgVecs<RV>& v = a.cg.v;
        if (first_it) {
            P.load(R, row, v.p);
            r = v.r[i];
            w = v.w[i];
            p = v.p[i];
            return;
        }
        X.load(R, row, v.x);
        x = v.x[i];
        if (upd) {
            U.load(R, row, a.u_in);
            P.load(R, row, v.p);
            r = v.r[i];
            w = v.w[i];
            s = v.s[i];
            u = a.u_in[i];
            p = v.p[i];
        }
    }
    // upd: this launch makes the update (otherwise x is final as it stands)
    __device__ __forceinline__ void apply(const Rows& R, const double* __restrict__ A64, const RV* __restrict__ D,
                                          const LastArgs<RV>& a, bool first_it, bool upd, RV al, RV be,
                                          double (&part)[3]) const
    {
        const CgVecs<RV>& v = a.cg.v;
        double ax = 0;
#pragma unroll
        for (int k = 0; k < R.width(); ++k) {
            RV xj;
            if (first_it) {
                xj = upd ? al * P.at(R, row, v.p, k) : RV(0);
            } else {
                const RV xo = X.at(R, row, v.x, k);
                xj = upd ? last_x(xo, U.at(R, row, a.u_in, k), P.at(R, row, v.p, k), al, be) : xo;
            }
            ax += A64[R.vidx(row, k)] * (double)xj;
        }
        const RV d = D[R.didx(row)];
        const RV xn = first_it ? (upd ? al * p : RV(0)) : (upd ? last_x(x, u, p, al, be) : x);
        const RV rt = (RV)(r0 - ax);
        const RV ut = d * rt;
        a.xf[i] = xn;
        a.v1.r[i] = rt;
        a.v1.p[i] = ut;
        a.v1.x[i] = RV(0);
        part[2] += (double)ut * (double)ut;
        if (upd) {
            const RV sn = first_it ? w : w + be * s;
            const RV un = d * (r - al * sn);
            part[1] += (double)un * (double)un;
        }
    }
};

// FIRST (it == 0) gathers p only and issues its first row with the
// reduction; later iterations gather x, u and p, after the reduction (held
// across it, three gathers would not fit the register budget).
template <class Rows, class RV, bool FIRST>
__global__ __launch_bounds__(kCgBlock) void cg_last_kernel(Rows R, MatVals mv, LastArgs<RV> a, int it)
{
    __shared__ double sA[kMaxPatterns * kMaxStencil];
    __shared__ RV sD[kMaxPatterns];
    const TableLoad<Rows, double> ta(mv.a64, R.value_count());
    const TableLoad<Rows, RV> td(dinv_vals<RV>(mv), R.diag_count());
    PartialLoad<3> pl(a.cg.pin, gridDim.x);
    SolveCtl* ctl = a.cg.ctl;
    const int poison = a.cg.clock->poison;
    const int done = ctl->done;
    const double tol2 = ctl->tol2;
    const int max_its = ctl->max_its;
    constexpr bool first_it = FIRST;
    const double2 h = a.cg.hist[first_it ? 0 : it - 1];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(a.cg.n, blk, gridDim.x);
    const int i0 = rr.lo + threadIdx.x;
    // every outcome reads the row (the update's inputs are a superset of the
    // final-x case's)
    LastRow<Rows, RV> q0;
    if (FIRST && i0 < rr.hi) q0.load(R, i0, a, true, true);
    const double* __restrict__ A64 = ta.store(sA, mv.a64);
    const RV* __restrict__ D = td.store(sD, dinv_vals<RV>(mv));
    double g[3];
    pl.total(g);    // its barriers also cover the table stores
    if (poison) return;
    const double gamma = g[0], delta = g[1], zeta = g[2];
    const bool lead = blk == 0 && threadIdx.x == 0;
    bool upd = false;
    double alpha = 0.0, beta = 0.0;
    if (done == 0) {
        if (!(zeta == zeta) || !(gamma == gamma) || !(delta == delta) || fabs(zeta) > DBL_MAX) {
            if (lead) { ctl->done = 3; ctl->it = it; ctl->zz = zeta; }
        } else if (zeta <= tol2) {
            if (lead) { ctl->done = 1; ctl->it = it; ctl->zz = zeta; }
        } else if (it >= max_its) {
            if (lead) { ctl->done = 2; ctl->it = it; ctl->zz = zeta; }
        } else {
            upd = true;
            if (first_it) {
                alpha = gamma / delta;
            } else {
                beta = gamma / h.x;
                alpha = gamma / (delta - beta * gamma / h.y);
            }
            if (lead) {
                a.cg.hist[it] = make_double2(gamma, alpha);
                ctl->it = it + 1;
                ctl->zz = zeta;
            }
        }
    }
    if (lead) {
        a.ctl1->done = 0;
        a.ctl1->it = 0;
    }
    const RV al = (RV)alpha, be = (RV)beta;
    double part[3] = {0.0, 0.0, 0.0};
    if (FIRST && i0 < rr.hi) q0.apply(R, A64, D, a, first_it, upd, al, be, part);
    for (int i = FIRST ? i0 + kCgBlock : i0; i < rr.hi; i += kCgBlock) {
        LastRow<Rows, RV> q;
        q.load(R, i, a, first_it, upd);
        q.apply(R, A64, D, a, first_it, upd, al, be, part);
    }
    write_partials<kCgBlock, 3>(part, a.cg.pout, blk);
}

// Start of a solve A x = b outside a step (the adjoint's solves), from x = 0
// or (warm) from the x the vector holds: stands in for the residual kernel,
// with r0 = b - A x, u0 = D^-1 r0 to p and the reductions cg_start_kernel
// reads (||D^-1 b||^2, 0, ||u0||^2); it also resets the solve's two control
// records.  A warm start gathers x, which then needs guard zones.
template <class Rows>
__global__ __launch_bounds__(kCgBlock) void plain_init_kernel(Rows R, MatVals mv, int n,
                                                              const double* __restrict__ b, CgVecs<double> v,
                                                              double* __restrict__ pout, SolveCtl* ctl, int max_its,
                                                              int warm)
{
    const double* __restrict__ D = mv.dinv64;
    const double* __restrict__ A = mv.a64;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    if (blk == 0 && threadIdx.x < 2) {
        SolveCtl& t = ctl[threadIdx.x];
        t.tol2 = 0.0;
        t.bb = 0.0;
        t.zz = 0.0;
        t.it = 0;
        t.done = 0;
        t.max_its = max_its;
        t.pad = 0;
    }
    const RowRange rr(n, blk, gridDim.x);
    double part[3] = {0.0, 0.0, 0.0};
    for (int i = rr.lo + threadIdx.x; i < rr.hi; i += kCgBlock) {
        const typename Rows::Row row = R.row(i);
        const double bi = b[i];
        const double d = D[R.didx(row)];
        double ri = bi;
        if (warm) {
            // r = b - A x from the x held (the previous solve's answer)
            Gathered<Rows, double> X;
            X.load(R, row, v.x);
            double ax = 0.0;
#pragma unroll
            for (int k = 0; k < R.width(); ++k) ax += A[R.vidx(row, k)] * X.at(R, row, v.x, k);
            ri = bi - ax;
        } else {
            v.x[i] = 0.0;
        }
        const double ui = d * ri;
        const double ub = d * bi;
        v.r[i] = ri;
        v.p[i] = ui;
        part[0] += ub * ub;
        part[2] += ui * ui;
    }
    write_partials<kCgBlock, 3>(part, pout, blk);
}

// Settles the last step of a batch (or a resumed step): the verdict on its
// final solve, V^{n+1} = Z + x1 (+ x2), the next initial guess and the
// activation times; advance_kernel then counts the step.  A solve still
// running poisons the clock until the host resumes it.  Launched only after a
// step, so that one is pending.
template <class RV>
struct FinishArgs {
    int n;
    StepClock* clock;
    SolveCtl* ctl_all;                 // both solves (statistics)
    int refined;
    int32_t* its_log;                  // [log_len][4]
    int log_len;
    double* __restrict__ V;
    SettleArgs<RV> s;
};

template <class RV>
__global__ __launch_bounds__(kCgBlock) void finish_kernel(FinishArgs<RV> a)
{
    StepClock* clock = a.clock;
    PartialLoad<3> pl(a.s.pin, a.s.gin);
    const int poison = clock->poison;
    const int k = clock->k;
    const int done = a.s.ctl->done;
    const double tol2 = a.s.ctl->tol2;
    const int done0 = (a.s.ctl0 ? a.s.ctl0 : a.s.ctl)->done;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(a.n, blk, gridDim.x);
    const int i0 = rr.lo + threadIdx.x, i1 = i0 + kCgBlock;
    RV d0 = 0, d1 = 0;
    double z0 = 0, z1 = 0;
    if (i0 < rr.hi) { d0 = a.s.x1[i0] + (a.s.x2 ? a.s.x2[i0] : RV(0)); z0 = a.s.Z[i0]; }
    if (i1 < rr.hi) { d1 = a.s.x1[i1] + (a.s.x2 ? a.s.x2[i1] : RV(0)); z1 = a.s.Z[i1]; }
    StepClock clk = *clock;
    // activation bookkeeping (V^k = Z - x0) and the next guess
    auto settle_guess = [&](int i, RV d, double z, double v) {
        const RV x0 = a.s.guess ? a.s.guess[i] : RV(0);
        if (a.s.act) track_activation(a.s, a.n, i, clk, k, z - (double)x0, v, step_time(clk, k));
        if (a.s.guess) a.s.guess[i] = next_guess(a.s, i, x0 + d, a.s.prev ? a.s.prev[i] : RV(0));
    };
    double g[3];
    pl.total(g);
    if (poison) return;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!solve_converged(a.s.ctl, done, tol2, g[2], lead)) {
        if (lead) {
            clock->poison = 1;
            clock->failed_step = k;
        }
        return;
    }
    if (lead && a.s.ctl0 && done0 == 0 && g[1] <= tol2) a.s.ctl0->done = 1;
    bool bad = false;
    if (i0 < rr.hi) {
        const double v = z0 + (double)d0;
        a.V[i0] = v;
        settle_guess(i0, d0, z0, v);
        bad |= !isfinite(v);
    }
    if (i1 < rr.hi) {
        const double v = z1 + (double)d1;
        a.V[i1] = v;
        settle_guess(i1, d1, z1, v);
        bad |= !isfinite(v);
    }
    for (int i = i1 + kCgBlock; i < rr.hi; i += kCgBlock) {
        const RV d = a.s.x1[i] + (a.s.x2 ? a.s.x2[i] : RV(0));
        const double z = a.s.Z[i];
        const double v = z + (double)d;
        a.V[i] = v;
        settle_guess(i, d, z, v);
        bad |= !isfinite(v);
    }
    if (bad) clock->nonfinite = 1;
}

// After a finish kernel: counts the step it settled (a poisoned clock means
// it did not).  A kernel of its own, so that no block of the finish kernel
// reads a clock field written in the same launch.
__global__ void advance_kernel(StepClock* clock, const SolveCtl* ctl_all, int refined, int32_t* its_log, int log_len)
{
    if (clock->poison || !clock->pending) return;
    log_step(its_log, log_len, clock->k, ctl_all, refined != 0);
    clock->k += 1;
    clock->pending = 0;
}

// Activation times reset to -1 and the voltage at t processed
// (AbstractCardiacProblem.cpp:459-464: the modifier sees the initial condition).
__global__ __launch_bounds__(kBlock) void activation_init_kernel(int n, const double* __restrict__ V,
                                                                 double* __restrict__ act, double thr, double t)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 5; ++c) act[(size_t)c * n + i] = -1.0;
    modifier_update(act, n, i, t, V[i], thr);
}

}  // namespace cmx

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
namespace cmx {

hipEvent_t take_event(cmx_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

const ModelOps* model_ops_lr91();
const ModelOps* model_ops_tt06();
const ModelOps* model_ops_ord();
const ModelOps* model_ops_hh52();
const ModelOps* model_ops_fox2002();
const ModelOps* model_ops_mahajan2008();
const ModelOps* model_ops_dn85();
const ModelOps* model_ops_fr2000();
const ModelOps* model_ops_maleckar2008();
const ModelOps* model_ops_n98();

// Plugin models (cmx_load_model), published once registered.
std::atomic<const ModelOps*> g_plugin_models[kMaxPluginModels];
std::mutex g_plugin_mutex;
int g_plugin_count = 0;

const ModelOps* model_ops(int model)
{
    if (model >= kPluginModelBase && model < kPluginModelBase + kMaxPluginModels)
        return g_plugin_models[model - kPluginModelBase].load(std::memory_order_acquire);
    switch (model) {
    case CMX_MODEL_LR91: return model_ops_lr91();
    case CMX_MODEL_TT06_EPI: return model_ops_tt06();
    case CMX_MODEL_ORD_ENDO: return model_ops_ord();
    case CMX_MODEL_HH52: return model_ops_hh52();
    case CMX_MODEL_FOX2002: return model_ops_fox2002();
    case CMX_MODEL_MAHAJAN2008: return model_ops_mahajan2008();
    case CMX_MODEL_DN85: return model_ops_dn85();
    case CMX_MODEL_FR2000: return model_ops_fr2000();
    case CMX_MODEL_MALECKAR2008: return model_ops_maleckar2008();
    case CMX_MODEL_N98: return model_ops_n98();
    default: return nullptr;
    }
}

}  // namespace cmx

using namespace cmx;

namespace {

thread_local std::string g_create_error;

int fail(cmx_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define CMX_HIP(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, CMX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <class T>
int dalloc(cmx_ctx* c, T** ptr, size_t count)
{
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, CMX_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    c->allocs.push_back(p);
    c->device_bytes += (double)bytes;
    *ptr = (T*)p;
    return CMX_OK;
}

// A vector the stencil may gather: c->pad zeroed guard elements on each side.
template <class T>
int dalloc_vec(cmx_ctx* c, T** ptr, size_t count)
{
    T* base = nullptr;
    const size_t total = count + 2 * (size_t)c->pad;
    int rc = dalloc(c, &base, total);
    if (rc) return rc;
    CMX_HIP(c, hipMemset(base, 0, total * sizeof(T)));
    *ptr = base + c->pad;
    return CMX_OK;
}

template <class T>
int upload(cmx_ctx* c, T** ptr, const T* host, size_t count)
{
    int rc = dalloc(c, ptr, count);
    if (rc) return rc;
    if (count) CMX_HIP(c, hipMemcpy(*ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
    return CMX_OK;
}

int model_sizes(int model, int* S, int* P)
{
    ModelMeta m;
    if (!model_meta(model, &m)) return CMX_E_ARG;
    *S = m.S;
    *P = m.P;
    return CMX_OK;
}

const double* model_init(int model) { ModelMeta m{}; model_meta(model, &m); return m.init; }
const double* model_pdef(int model) { ModelMeta m{}; model_meta(model, &m); return m.pdef; }

// A (model, integrator) pair the cell kernels are built for.
bool model_supports(int model, int integrator)
{
    ModelMeta m;
    if (!model_meta(model, &m)) return false;
    return integrator == CMX_FORWARD_EULER || integrator == CMX_RUSH_LARSEN ||
           (m.implicit && integrator >= CMX_BACKWARD_EULER && integrator <= CMX_GRL2);
}

int grid_of(int n) { return (n + kBlock - 1) / kBlock; }

constexpr int kMaxGraphSteps = 32;    // cap on the steps of the largest captured graph
constexpr size_t kGraphCache = 64;    // instantiated graphs kept per context

// Destroys every captured graph once the stream has drained (a graph still
// executing must not be destroyed).
void drop_graphs(cmx_ctx* c)
{
    if (c->graphs.empty()) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    c->graphs.clear();
}

// rocprofv3 preloads its tool library and exports ROCPROF_* settings; under
// its dispatch interception the context runs plain launches (graph replay
// under the tracer is not used).
bool profiler_attached()
{
    const char* pre = std::getenv("LD_PRELOAD");
    if (pre && std::strstr(pre, "rocprof")) return true;
    return std::getenv("ROCPROF_OUTPUT_PATH") || std::getenv("ROCPROFILER_LIBRARY_CTOR");
}

// Folds completed event pairs into the totals (call after a stream sync).
void collect_times(cmx_ctx* c)
{
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->k_ms[p.kind] += ms;
            c->k_launches[p.kind] += 1;
        }
        c->event_pool.push_back(p.a);
        c->event_pool.push_back(p.b);
    }
    c->pending.clear();
}

MatVals mat_vals(cmx_ctx* c)
{
    MatVals m;
    m.m64 = c->m64;
    m.k64 = c->k64;
    m.a64 = c->a64;
    m.a32 = c->a32;
    m.dinv64 = c->dinv64;
    m.dinv32 = c->dinv32;
    return m;
}

// Partial-sum buffers: kResPart for the residual kernels, kIterPart + (it & 1)
// for the reduction that feeds iteration it of a solve, kFoldPart for the
// folded last iteration (read by the settling kernel, and by the refinement
// solve's start when the host resumes it).
constexpr int kIterPart = 0;
constexpr int kResPart = 2;
constexpr int kFoldPart = 3;
constexpr int kPartBufs = 4;
double* pbuf(cmx_ctx* c, int k) { return c->partials + (size_t)k * kMaxBlocks * 3; }

// Vector slots of cmx_ctx::vec.
// V_R2 / V_P2: the refinement solve's r and p; V_XF: the first solve's answer
// and V_U: D^-1 r for the last iteration when the refinement is folded into it.
enum { V_R = 0, V_W, V_Z, V_S, V_P, V_MA, V_MB, V_X1, V_X2, V_X0, V_XP, V_R2, V_P2, V_XF, V_U, V_COUNT };

template <class RV>
CgVecs<RV> cg_vecs(cmx_ctx* c, int solve, int it)
{
    CgVecs<RV> v;
    v.r = (RV*)c->vec[solve ? V_R2 : V_R];
    v.w = (RV*)c->vec[V_W];
    v.z = (RV*)c->vec[V_Z];
    v.s = (RV*)c->vec[V_S];
    v.p = (RV*)c->vec[solve ? V_P2 : V_P];
    // iteration it reads the m written by iteration it-1 (the start kernel
    // writes it for iteration 0)
    v.m_in = (RV*)c->vec[(it & 1) ? V_MB : V_MA];
    v.m_out = (RV*)c->vec[(it & 1) ? V_MA : V_MB];
    v.x = (RV*)c->vec[solve ? V_X2 : V_X1];
    return v;
}

template <class RV>
CgArgs<RV> cg_args(cmx_ctx* c, int solve)
{
    CgArgs<RV> a;
    a.n = c->n;
    a.first = 0;
    a.rtol2 = c->rtol * c->rtol;
    a.atol2 = c->atol * c->atol;
    a.clock = c->clock;
    a.ctl = c->ctl + solve;
    a.ctl_all = c->ctl;
    a.hist = c->hist;
    a.pin = pbuf(c, kResPart);
    a.pout = pbuf(c, kIterPart);
    a.v = cg_vecs<RV>(c, solve, 0);
    return a;
}

bool refined_of(const cmx_ctx* c) { return c->precision == CMX_MIXED && c->refine > 0; }

// Steps at the current budgets fold the refinement residual into the first
// solve's last iteration (cg_last_kernel): refined, the refinement solve
// without a budget of its own, and at least one iteration budgeted.
bool fused_of(const cmx_ctx* c) { return refined_of(c) && c->fold_refine && c->refine_budget == 0 && c->budget >= 1; }

// The reduction that decides a step's final solve at the current budgets: the
// last iteration's (or the start kernel's when the budget is 0), or the
// refinement residual's when the refinement solve has no budget -- written by
// the last iteration itself when the refinement is folded in.
const double* verdict_buf(cmx_ctx* c)
{
    if (fused_of(c)) return pbuf(c, kFoldPart);
    if (refined_of(c))
        return c->refine_budget > 0 ? pbuf(c, kIterPart + (c->refine_budget & 1)) : pbuf(c, kResPart);
    return pbuf(c, kIterPart + (c->budget & 1));
}

// fused: the pending step folded its refinement residual into the first
// solve's last iteration (its answer is in V_XF); rec0: the verdict is that
// iteration's reduction, whose slot 1 tells whether the first solve converged.
template <class RV>
SettleArgs<RV> settle_args(cmx_ctx* c, const double* verdict, bool fused, bool rec0)
{
    const bool refined = refined_of(c);
    SettleArgs<RV> s;
    s.ctl = c->ctl + (refined ? 1 : 0);
    s.ctl0 = fused && rec0 ? c->ctl : nullptr;
    s.pin = verdict;
    s.gin = c->G;
    s.Z = c->extrapolate ? c->Y : c->V;
    s.x1 = (const RV*)c->vec[fused ? V_XF : V_X1];
    s.x2 = refined ? (const RV*)c->vec[V_X2] : nullptr;
    s.guess = c->extrapolate ? (RV*)c->vec[V_X0] : nullptr;
    s.prev = c->extrapolate >= 2 ? (RV*)c->vec[V_XP] : nullptr;
    s.act = c->act_every > 0 ? c->act : nullptr;
    s.act_thr = c->act_thr;
    s.act_every = c->act_every > 0 ? c->act_every : 1;
    return s;
}

// ----- typed launches -------------------------------------------------------
template <class RS, class RV>
void launch_cell_model(cmx_ctx* c, int model, int integrator, int mode, int kind, int G, const CellArgs<RS, RV>& a)
{
    cell_launch<RS, RV>(model_ops(model))(c, integrator, mode, kind, G, a);
}

// The cell kernel in one of its roles; verdict: the reduction that decides
// the pending step's final solve.
template <class RS, class RV>
int launch_cells(cmx_ctx* c, int mode, const double* verdict, int kind, bool fused, bool rec0)
{
    CellArgs<RS, RV> a;
    a.n = c->n;
    a.n_cells = c->n;
    a.node = nullptr;
    a.n_sub = c->n_sub;
    a.n_half = c->n_half;
    a.dt_ode = c->dt_ode;
    a.am = c->am;
    a.cm = c->cm;
    a.clock = c->clock;
    a.V = c->V;
    a.ps = c->ps;
    a.state = (RS*)c->state;
    a.params = (const RS*)c->params;
    a.pdef = (const RS*)c->pdef;
    a.n_param_planes = c->n_param_planes;
    a.stim_id = c->stim_id;
    a.stim = c->stim;
    a.F = (RV*)c->F;
    a.iion_out = (c->svi && mode == kModeIci) ? (RV*)c->svi_iion : nullptr;
    a.s = settle_args<RV>(c, verdict, fused, rec0);
    a.c = c->am * c->cm / c->dt;
    a.Y = c->Y;
    a.G1 = (RV*)c->G1;
    a.trace = mode == kModeIci ? c->trace : nullptr;
    a.trace_k0 = c->trace_k0;
    a.trace_len = c->trace_len;
    a.trace_slot = c->trace_slot;
    a.lut = (std::is_same<RS, float>::value && c->cell_groups.empty()) ? c->lut : nullptr;
    a.lut_vmin = (float)c->lut_vmin;
    a.lut_inv_dv = c->lut_dv > 0 ? (float)(1.0 / c->lut_dv) : 0.0f;
    a.lut_rows = c->lut_rows;
    if (c->cell_groups.empty()) {
        launch_cell_model<RS, RV>(c, c->model, c->integrator, mode, kind, grid_of(c->n), a);
    } else {
        // one launch per cell type; each settles the pending step at its own
        // nodes (a failed verdict poisons the clock in the first launch and
        // the others return at once)
        for (const CellGroup& g : c->cell_groups) {
            CellArgs<RS, RV> b = a;
            b.n_cells = g.n;
            b.node = g.node;
            b.ps = g.ps;
            b.state = (RS*)g.state;
            b.params = nullptr;
            b.pdef = (const RS*)g.pdef;
            b.n_param_planes = 0;
            b.trace = nullptr;
            launch_cell_model<RS, RV>(c, g.model, g.integrator, mode, kind, grid_of(g.n), b);
        }
    }
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

template <class Rows, class RS, class RV>
int launch_svi(cmx_ctx* c, const Rows& R)
{
    SviArgs<RS, RV> a;
    a.n = c->n;
    a.nb = c->dim + 1;
    a.nq = c->nq;
    a.am = c->am;
    a.cm = c->cm;
    a.clock = c->clock;
    a.elems = c->elems;
    a.edet = c->edet;
    a.inc_ptr = c->inc_ptr;
    a.inc = c->inc;
    a.iion = (const RV*)c->svi_iion;
    a.V = c->V;
    a.state = (const RS*)c->state;
    a.ps = c->ps;
    a.params = (const RS*)c->params;
    a.pdef = (const RS*)c->pdef;
    a.n_param_planes = c->n_param_planes;
    for (int q = 0; q < 4; ++q) {
        a.qw[q] = c->qw[q];
        for (int k = 0; k < 3; ++k) a.qp[q][k] = c->qp[q][k];
    }
    a.out = (RV*)c->svi_out;
    svi_launch<RS, RV>(model_ops(c->model), RowsSlot<Rows>::k)(c, &R, a, grid_of(c->n));
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

template <class Rows, class RV>
int launch_residual(cmx_ctx* c, const Rows& R, int solve)
{
    ResidualArgs<RV> a;
    a.n = c->n;
    a.c = c->am * c->cm / c->dt;
    a.clock = c->clock;
    a.its_log = c->its_log;
    a.log_len = c->batch;
    a.refined = refined_of(c) ? 1 : 0;
    a.Z = c->extrapolate ? c->Y : c->V;
    a.G = (const RV*)(c->extrapolate ? c->G1 : c->F);
    a.svi = (c->svi && !c->split) ? (const RV*)c->svi_out : nullptr;
    a.r0_64 = c->r0_64;
    a.x1 = (const RV*)c->vec[V_X1];
    a.pin0 = pbuf(c, kIterPart + (c->budget & 1));
    a.gin0 = c->G;
    a.ctl0 = c->ctl;
    a.v = cg_vecs<RV>(c, solve, 0);
    a.partials = pbuf(c, kResPart);
    a.ctl = c->ctl + solve;
    if (solve == 1)
        CMX_LAUNCH(c, CMX_K_RESIDUAL, (refine_residual_kernel<Rows, RV>), c->G, kCgBlock, R, mat_vals(c), a);
    else
        CMX_LAUNCH(c, CMX_K_RESIDUAL, (residual_kernel<Rows, RV>), c->G, kCgBlock, R, mat_vals(c), a);
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

// Rows per thread the CG kernels hold in registers: the context's knob, or
// the precision's default (2 in FP32, 1 in FP64).
template <class RV>
int cg_rows_of(const cmx_ctx* c)
{
    return c->cg_rows > 0 ? c->cg_rows : (sizeof(RV) == 4 ? 2 : 1);
}

// Resident blocks per CU of the CG iteration kernel the context will launch.
template <class Rows, class RV>
int iter_blocks_per_cu(const cmx_ctx* c)
{
    int b = 0;
    const hipError_t e = cg_rows_of<RV>(c) == 2
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, cg_iter_kernel<Rows, RV, 2>, kCgBlock, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, cg_iter_kernel<Rows, RV, 1>, kCgBlock, 0);
    return e == hipSuccess ? std::max(b, 1) : 1;
}

template <class RV>
int iter_blocks_per_cu_rv(const cmx_ctx* c)
{
    if (c->format != CMX_FMT_STENCIL) return iter_blocks_per_cu<EllRows, RV>(c);
    if (c->width == 4) return iter_blocks_per_cu<StencilRows<4>, RV>(c);
    if (c->width == 8) return iter_blocks_per_cu<StencilRows<8>, RV>(c);
    return iter_blocks_per_cu<StencilRows<16>, RV>(c);
}

// Default solver blocks: one per CU, or on meshes with at least 8 rows per
// thread at that count, as many as the CG iteration kernel keeps resident
// (each thread walks its rows one after another, so the resident wavefronts
// bound the loads in flight; the FP32 ELL iteration on a 13M-row shell union
// runs 1.5x faster at two blocks per CU).  cmx_autotune and
// cmx_set_launch_config override it.
int default_solver_blocks(const cmx_ctx* c)
{
    const int cg_rows = (c->n + kCgBlock - 1) / kCgBlock;
    int G = std::min(cg_rows, kCgGrid);
    if ((long)c->n >= 8L * kCgGrid * kCgBlock) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            return G;
        const int per_cu = c->precision == CMX_MIXED ? iter_blocks_per_cu_rv<float>(c) : iter_blocks_per_cu_rv<double>(c);
        G = std::min({cg_rows, per_cu * cus, kMaxBlocks});
    }
    return G;
}

// Start of a solve (w0 = A u0 and its dot products; ends a solve whose
// residual is already below the threshold).
template <class Rows, class RV>
int launch_cg_start(cmx_ctx* c, const Rows& R, int solve, const double* pin = nullptr)
{
    CgArgs<RV> a = cg_args<RV>(c, solve);
    a.first = solve == 0 ? 1 : 0;
    if (pin) a.pin = pin;   // the reduction holding ||D^-1 r0||^2 of this solve
    a.v.m_out = (RV*)c->vec[V_MA];   // read by iteration 0
    if (cg_rows_of<RV>(c) == 2) CMX_LAUNCH(c, CMX_K_SPMV, (cg_start_kernel<Rows, RV, 2>), c->G, kCgBlock, R, mat_vals(c), a);
    else CMX_LAUNCH(c, CMX_K_SPMV, (cg_start_kernel<Rows, RV, 1>), c->G, kCgBlock, R, mat_vals(c), a);
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

// CG iterations from .. from+count-1 of a solve, one kernel each; u_last:
// the last of them also stores D^-1 r for a folded last iteration.
template <class Rows, class RV>
int launch_cg_iters(cmx_ctx* c, const Rows& R, int solve, int from, int count, bool u_last = false)
{
    CgArgs<RV> a = cg_args<RV>(c, solve);
    const MatVals mv = mat_vals(c);
    for (int it = from; it < from + count; ++it) {
        a.pin = pbuf(c, kIterPart + (it & 1));
        a.pout = pbuf(c, kIterPart + ((it + 1) & 1));
        a.v = cg_vecs<RV>(c, solve, it);
        if (u_last && it == from + count - 1) a.v.u_out = (RV*)c->vec[V_U];
        if (cg_rows_of<RV>(c) == 2) CMX_LAUNCH(c, CMX_K_UPDATE, (cg_iter_kernel<Rows, RV, 2>), c->G, kCgBlock, R, mv, a, it);
        else CMX_LAUNCH(c, CMX_K_UPDATE, (cg_iter_kernel<Rows, RV, 1>), c->G, kCgBlock, R, mv, a, it);
    }
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

// Iteration `it` of the first solve as its last, with the refinement
// residual folded in (FP32 vectors only).
template <class Rows, class RV>
int launch_cg_last(cmx_ctx* c, const Rows& R, int it)
{
    if constexpr (std::is_same<RV, float>::value) {
        LastArgs<RV> a;
        a.cg = cg_args<RV>(c, 0);
        a.cg.pin = pbuf(c, kIterPart + (it & 1));
        a.cg.pout = pbuf(c, kFoldPart);
        a.cg.v = cg_vecs<RV>(c, 0, it);
        a.r0_64 = c->r0_64;
        a.u_in = (const RV*)c->vec[V_U];
        a.xf = (RV*)c->vec[V_XF];
        a.ctl1 = c->ctl + 1;
        a.v1 = cg_vecs<RV>(c, 1, 0);
        if (it == 0) CMX_LAUNCH(c, CMX_K_UPDATE, (cg_last_kernel<Rows, RV, true>), c->G, kCgBlock, R, mat_vals(c), a, it);
        else CMX_LAUNCH(c, CMX_K_UPDATE, (cg_last_kernel<Rows, RV, false>), c->G, kCgBlock, R, mat_vals(c), a, it);
        CMX_HIP(c, hipGetLastError());
        return CMX_OK;
    } else {
        return fail(c, CMX_E_STATE, "the folded refinement runs on FP32 vectors only");
    }
}

// Settles the pending step; verdict: the reduction that decides its final
// solve.
template <class RS, class RV>
int launch_finish(cmx_ctx* c, const double* verdict, bool fused, bool rec0)
{
    if (c->split) {
        // operator splitting: the last half-step of the cells completes the step
        int rc = launch_cells<RS, RV>(c, kModeSplitFinal, verdict, CMX_K_FINISH, fused, rec0);
        if (rc) return rc;
        advance_kernel<<<1, 1, 0, c->stream>>>(c->clock, c->ctl, refined_of(c) ? 1 : 0, c->its_log, c->batch);
        CMX_HIP(c, hipGetLastError());
        return CMX_OK;
    }
    FinishArgs<RV> a;
    a.n = c->n;
    a.clock = c->clock;
    a.ctl_all = c->ctl;
    a.refined = refined_of(c) ? 1 : 0;
    a.its_log = c->its_log;
    a.log_len = c->batch;
    a.V = c->V;
    a.s = settle_args<RV>(c, verdict, fused, rec0);
    CMX_LAUNCH(c, CMX_K_FINISH, (finish_kernel<RV>), c->G, kCgBlock, a);
    advance_kernel<<<1, 1, 0, c->stream>>>(c->clock, c->ctl, a.refined, c->its_log, c->batch);
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

// Enqueue one whole step with iteration budgets B (first solve) and B2
// (refinement solve): cells (which settle the previous step), residual,
// start, B iterations, [refinement residual, (start, B2 iterations)].  With
// B >= 1 and B2 = 0 the refinement residual is folded into the B-th
// iteration.  The step stays pending until the next step's cell kernel or the
// batch's finish kernel settles it.  A solve that needs more than its budget
// is either continued by the refinement solve (first solve) or resumed by the
// host after the settling kernel poisons the clock.
template <class Rows, class RS, class RV>
int enqueue_step(cmx_ctx* c, const Rows& R)
{
    int rc;
    const bool fused = fused_of(c);
    if ((rc = launch_cells<RS, RV>(c, c->split ? kModeSplit : kModeIci, verdict_buf(c), CMX_K_CELL, fused, fused)))
        return rc;
    if (c->svi && !c->split && (rc = launch_svi<Rows, RS, RV>(c, R))) return rc;
    if ((rc = launch_residual<Rows, RV>(c, R, 0))) return rc;
    if ((rc = launch_cg_start<Rows, RV>(c, R, 0))) return rc;
    if (fused) {
        if ((rc = launch_cg_iters<Rows, RV>(c, R, 0, 0, c->budget - 1, true))) return rc;
        return launch_cg_last<Rows, RV>(c, R, c->budget - 1);
    }
    if ((rc = launch_cg_iters<Rows, RV>(c, R, 0, 0, c->budget))) return rc;
    if (refined_of(c)) {
        // FP64 refinement: true residual of V^n + x1 against the step's b,
        // then a second CG solve on it into x2 when the budget allows one.
        if ((rc = launch_residual<Rows, RV>(c, R, 1))) return rc;
        if (c->refine_budget > 0) {
            if ((rc = launch_cg_start<Rows, RV>(c, R, 1))) return rc;
            if ((rc = launch_cg_iters<Rows, RV>(c, R, 1, 0, c->refine_budget))) return rc;
        }
    }
    return CMX_OK;
}

int sync_state(cmx_ctx* c)
{
    CMX_HIP(c, hipMemcpyAsync(c->h_clock, c->clock, sizeof(StepClock), hipMemcpyDeviceToHost, c->stream));
    CMX_HIP(c, hipMemcpyAsync(c->h_ctl, c->ctl, 2 * sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->pending.empty()) collect_times(c);
    return CMX_OK;
}

// Instantiated graph of `steps` consecutive steps at the current budgets
// (captured once; kernel arguments are fixed because time lives in the
// device step clock).  Returns nullptr when capture is unavailable.  The
// cache is bounded; an evicted graph may still be executing on the stream
// (it can be the one launched just before this capture), so the stream is
// drained before the executable is destroyed.
template <class Rows, class RS, class RV>
hipGraphExec_t step_graph(cmx_ctx* c, const Rows& R, int steps)
{
    for (auto& g : c->graphs)
        if (g.budget == c->budget && g.refine_budget == c->refine_budget && g.steps == steps && g.G == c->G &&
            g.rows == c->cg_rows && g.act_every == c->act_every && g.act_thr == c->act_thr)
            return g.exec;
    if (c->graphs.size() >= kGraphCache) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        (void)hipGraphExecDestroy(c->graphs.front().exec);
        (void)hipGraphDestroy(c->graphs.front().graph);
        c->graphs.erase(c->graphs.begin());
        ++c->graph_evictions;
    }
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    int rc = CMX_OK;
    for (int s = 0; s < steps && rc == CMX_OK; ++s) rc = enqueue_step<Rows, RS, RV>(c, R);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != CMX_OK || e != hipSuccess || !graph) {
        (void)hipGetLastError();
        if (graph) (void)hipGraphDestroy(graph);
        return nullptr;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipGraphDestroy(graph);
        return nullptr;
    }
    // the executable's one-time device set-up now, on the stream, rather
    // than in its first replay (a primed graph's first replay is often timed)
    if (hipGraphUpload(exec, c->stream) != hipSuccess) (void)hipGetLastError();
    c->graphs.push_back({c->budget, c->refine_budget, steps, c->G, c->cg_rows, c->act_every, c->act_thr, graph, exec});
    ++c->graph_captures;
    return exec;
}

bool graphs_on(const cmx_ctx* c) { return c->use_graph && !c->graph_profiler_off && c->prof_mask == 0; }

// Enqueues `todo` steps: replays of captured graphs of graph_steps, then
// halving numbers of steps down to 1, when enabled (and no kernel is being timed), plain launches otherwise.
template <class Rows, class RS, class RV>
int enqueue_steps(cmx_ctx* c, const Rows& R, int todo)
{
    if (graphs_on(c)) {
        int left = todo;
        while (left > 0) {
            int g = c->graph_steps;
            while (g > left) g >>= 1;
            hipGraphExec_t exec = step_graph<Rows, RS, RV>(c, R, g);
            if (!exec) break;
            CMX_HIP(c, hipGraphLaunch(exec, c->stream));
            ++c->graph_launches;
            left -= g;
        }
        todo = left;
    }
    for (int s = 0; s < todo; ++s) {
        const int rc = enqueue_step<Rows, RS, RV>(c, R);
        if (rc) return rc;
    }
    return CMX_OK;
}

// Captures the graphs a batch of `todo` steps replays at the current budgets
// (outside any timed region: cmx_autotune calls it for the next batch).
template <class Rows, class RS, class RV>
int prime_graphs(cmx_ctx* c, const Rows& R, int todo)
{
    if (!graphs_on(c)) return CMX_OK;
    for (int g = c->graph_steps; g >= 1; g >>= 1)
        if (todo >= g && !step_graph<Rows, RS, RV>(c, R, g)) break;
    return CMX_OK;
}

// A cell-kernel failure flagged in the clock (a root cause that can also make
// the following solve fail, so it is reported first), or CMX_OK.
int cell_failure(cmx_ctx* c)
{
    if (c->h_clock->nonfinite == 2)
        return fail(c, CMX_E_NONFINITE, "backward-Euler cell: Newton iteration did not converge in 16 steps");
    if (c->h_clock->nonfinite == 4)
        return fail(c, CMX_E_NONFINITE, "membrane voltage outside the gate lookup table (-250..350 mV)");
    return CMX_OK;
}

// Runs the clock's steps k .. n_steps-1 in batches, resuming poisoned steps.
template <class Rows, class RS, class RV>
int run_steps(cmx_ctx* c, const Rows& R)
{
    const bool refined = c->precision == CMX_MIXED && c->refine > 0;
    const int solve = refined ? 1 : 0;
    while (c->h_clock->k < c->h_clock->n_steps) {
        const int k_start = c->h_clock->k;
        const int todo = std::min(c->h_clock->n_steps - k_start, c->batch);
        const bool fused = fused_of(c);   // at the budgets the batch runs with
        int rc = enqueue_steps<Rows, RS, RV>(c, R, todo);
        if (rc) return rc;
        if ((rc = launch_finish<RS, RV>(c, verdict_buf(c), fused, fused))) return rc;
        rc = sync_state(c);
        if (rc) return rc;
        if ((rc = cell_failure(c))) return rc;
        int guard = 0;
        while (c->h_clock->poison) {
            if (++guard > c->max_its + 16) return fail(c, CMX_E_STATE, "resuming a step made no progress");
            // Resume the failed step's final solve where it stopped.
            const SolveCtl& ctl = c->h_ctl[solve];
            if (ctl.done == 2)
                return fail(c, CMX_E_DIVERGED, "CG reached the iteration limit (" + std::to_string(c->max_its) +
                                                   ") without converging: KSP_DIVERGED_ITS");
            if (ctl.done == 3) return fail(c, CMX_E_NONFINITE, "non-finite residual in the diffusion solve");
            ++c->resumes;
            if (solve == 1) c->refine_budget = std::min(c->refine_budget + 2, c->max_its);
            else c->budget = std::min(c->budget + 2, c->max_its);
            // Clear the poison first: every kernel launched while it is set is a no-op.
            CMX_HIP(c, hipMemsetAsync(&c->clock->poison, 0, sizeof(int32_t), c->stream));
            const double* verdict = pbuf(c, kResPart);
            if (ctl.done == 0) {
                const int from = ctl.it;
                // the start kernel is idempotent; a refinement solve with a
                // zero budget has not run it yet
                // (after a folded step its residual's norm is in the folded reduction)
                if (from == 0 && (rc = launch_cg_start<Rows, RV>(c, R, solve, fused ? pbuf(c, kFoldPart) : nullptr)))
                    return rc;
                const int chunk = std::max(1, std::min(4, c->max_its + 1 - from));
                if ((rc = launch_cg_iters<Rows, RV>(c, R, solve, from, chunk))) return rc;
                verdict = pbuf(c, kIterPart + ((from + chunk) & 1));
            }
            if ((rc = launch_finish<RS, RV>(c, verdict, fused, false))) return rc;
            if ((rc = sync_state(c))) return rc;
        }
        const int got = c->h_clock->k - k_start;
        if (got <= 0) return fail(c, CMX_E_STATE, "no step completed in a batch");
        // iteration statistics and the next budgets
        CMX_HIP(c, hipMemcpy(c->h_its, c->its_log, 4 * c->batch * sizeof(int32_t), hipMemcpyDeviceToHost));
        int mx0 = 0, mx1 = 0;
        bool short0 = false;
        for (int s = 0; s < got; ++s) {
            const int32_t* slot = c->h_its + 4 * ((k_start + s) % c->batch);
            c->cg_total += slot[0] + slot[2];
            mx0 = std::max(mx0, slot[0]);
            mx1 = std::max(mx1, slot[2]);
            short0 |= slot[1] == 0;
            c->last_its = slot[0] + slot[2];
        }
        if (refined) c->refine_total += got;
        if (fused) c->folded_total += got;
        // budgets: the largest count of the batch; a first solve that the
        // refinement had to finish gets one more iteration
        c->budget = std::min(short0 ? std::max(mx0, c->budget) + 1 : mx0, c->max_its);
        c->refine_budget = std::min(mx1, c->max_its);
        c->steps += got;
        if ((rc = cell_failure(c))) return rc;
        if (c->h_clock->nonfinite)
            return fail(c, CMX_E_NONFINITE, "non-finite transmembrane potential (cell ODE or solve blew up)");
    }
    c->last_residual = std::sqrt(std::max(c->h_ctl[solve].zz, 0.0));
    return CMX_OK;
}

template <int W>
StencilRows<W> stencil_rows(cmx_ctx* c)
{
    StencilRows<W> R;
    R.pid = c->pid;
    R.n = c->n;
    R.n_patterns = c->n_patterns;
    R.pad = c->pad;
    for (int k = 0; k < W; ++k) R.off[k] = c->h_off[k];
    return R;
}

template <class RS, class RV>
int steps_typed(cmx_ctx* c)
{
    if (c->format == CMX_FMT_STENCIL) {
        if (c->width == 4) return run_steps<StencilRows<4>, RS, RV>(c, stencil_rows<4>(c));
        if (c->width == 8) return run_steps<StencilRows<8>, RS, RV>(c, stencil_rows<8>(c));
        return run_steps<StencilRows<16>, RS, RV>(c, stencil_rows<16>(c));
    }
    return run_steps<EllRows, RS, RV>(c, ell_rows(c));
}

template <class RS, class RV>
int prime_typed(cmx_ctx* c, int todo)
{
    if (c->format == CMX_FMT_STENCIL) {
        if (c->width == 4) return prime_graphs<StencilRows<4>, RS, RV>(c, stencil_rows<4>(c), todo);
        if (c->width == 8) return prime_graphs<StencilRows<8>, RS, RV>(c, stencil_rows<8>(c), todo);
        return prime_graphs<StencilRows<16>, RS, RV>(c, stencil_rows<16>(c), todo);
    }
    return prime_graphs<EllRows, RS, RV>(c, ell_rows(c), todo);
}

template <class Rows>
CgArgs<double> plain_args(cmx_ctx* c, const PlainCg& s, int it)
{
    CgArgs<double> a;
    a.n = c->n;
    a.first = 0;
    a.rtol2 = s.rtol * s.rtol;
    a.atol2 = 0.0;
    a.clock = c->clock;
    a.ctl = s.ctl;
    a.ctl_all = s.ctl;
    a.hist = s.hist;
    a.pin = s.part[it & 1];
    a.pout = s.part[(it + 1) & 1];
    a.v.r = s.r;
    a.v.w = s.w;
    a.v.z = s.z;
    a.v.s = s.s;
    a.v.p = s.p;
    a.v.m_in = s.m[it & 1];
    a.v.m_out = s.m[(it + 1) & 1];
    a.v.x = s.x;
    return a;
}

template <class Rows>
int plain_start_rows(cmx_ctx* c, const Rows& R, const PlainCg& s, const double* b)
{
    CgArgs<double> a = plain_args<Rows>(c, s, 0);
    plain_init_kernel<Rows><<<c->G, kCgBlock, 0, c->stream>>>(R, mat_vals(c), c->n, b, a.v, s.part[2], s.ctl,
                                                              s.max_its, s.warm);
    a.first = 1;
    a.pin = s.part[2];
    a.pout = s.part[0];
    a.v.m_out = s.m[0];   // read by iteration 0
    cg_start_kernel<Rows, double><<<c->G, kCgBlock, 0, c->stream>>>(R, mat_vals(c), a);
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

template <class Rows>
int plain_iters_rows(cmx_ctx* c, const Rows& R, const PlainCg& s, int from, int count)
{
    const MatVals mv = mat_vals(c);
    for (int it = from; it < from + count; ++it) {
        const CgArgs<double> a = plain_args<Rows>(c, s, it);
        cg_iter_kernel<Rows, double><<<c->G, kCgBlock, 0, c->stream>>>(R, mv, a, it);
    }
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

// Host [planes][n] doubles <-> device [planes][stride] RS (stride >= n; the
// gap elements stay as they are on the device).
template <class RS>
int to_device_planes(cmx_ctx* c, void* dst, const double* src, size_t planes, size_t n, size_t stride)
{
    std::vector<RS> tmp(planes * stride, RS(0));
    for (size_t q = 0; q < planes; ++q) std::copy(src + q * n, src + (q + 1) * n, tmp.begin() + q * stride);
    CMX_HIP(c, hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(RS), hipMemcpyHostToDevice));
    return CMX_OK;
}

template <class RS>
int from_device_planes(cmx_ctx* c, double* dst, const void* src, size_t planes, size_t n, size_t stride)
{
    std::vector<RS> tmp(planes * stride);
    CMX_HIP(c, hipMemcpy(tmp.data(), src, tmp.size() * sizeof(RS), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < planes; ++q) std::copy(tmp.begin() + q * stride, tmp.begin() + q * stride + n, dst + q * n);
    return CMX_OK;
}

double host_stimulus(const cmx_stimulus& s, double t)
{
    if (s.period > 0.0) {
        if (t < s.start) return 0.0;
        const double beat = std::fmod(t - s.start, s.period);
        return (beat >= 0.0 && beat <= s.duration && t <= s.stop) ? s.magnitude : 0.0;
    }
    return (s.start <= t && t <= s.duration + s.start) ? s.magnitude : 0.0;
}

// State-variable interpolation tables: element connectivity, |det J| per
// element, each node's incident (element, local index) list in ascending
// element order, and the second-order quadrature rule of the element
// dimension (GaussianQuadratureRule(2), pde/src/solver/GaussianQuadratureRule.cpp:
// 81-91, 126-135, 184-202).
template <class VecAlloc>
int setup_svi(cmx_ctx* c, const cmx_mesh* mesh, VecAlloc& rvec)
{
    int rc;
    const int n = c->n, dim = c->dim, nb = dim + 1, ne = mesh->n_elems;
    if ((rc = rvec(&c->svi_iion))) return rc;
    if ((rc = rvec(&c->svi_out))) return rc;
    std::vector<int32_t> ptr(n + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < nb; ++k) ptr[mesh->elems[(size_t)e * nb + k] + 1] += 1;
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> inc(ptr[n]), fill(ptr.begin(), ptr.end() - 1);
    std::vector<double> det(ne);
    for (int e = 0; e < ne; ++e) {
        const int32_t* nd = mesh->elems + (size_t)e * nb;
        for (int k = 0; k < nb; ++k) inc[fill[nd[k]]++] = 4 * e + k;
        double J[3][3] = {};
        for (int r = 0; r < dim; ++r)
            for (int q = 0; q < dim; ++q)
                J[r][q] = mesh->coords[(size_t)nd[q + 1] * dim + r] - mesh->coords[(size_t)nd[0] * dim + r];
        double d;
        if (dim == 1) d = J[0][0];
        else if (dim == 2) d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else
            d = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        det[e] = std::fabs(d);
    }
    if ((rc = upload(c, &c->elems, mesh->elems, (size_t)ne * nb))) return rc;
    if ((rc = upload(c, &c->edet, det.data(), det.size()))) return rc;
    if ((rc = upload(c, &c->inc_ptr, ptr.data(), ptr.size()))) return rc;
    if ((rc = upload(c, &c->inc, inc.data(), inc.size()))) return rc;
    if (dim == 1) {
        const double s = std::sqrt(1.0 / 3.0);
        c->nq = 2;
        c->qw[0] = c->qw[1] = 0.5;
        c->qp[0][0] = (-s + 1.0) / 2.0;
        c->qp[1][0] = (s + 1.0) / 2.0;
    } else if (dim == 2) {
        c->nq = 3;
        c->qw[0] = c->qw[1] = c->qw[2] = 1.0 / 6.0;
        const double pts[3][2] = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < 2; ++k) c->qp[q][k] = pts[q][k];
    } else {
        const double s = std::sqrt(0.2);
        const double a = (1.0 + 3.0 * s) / 4.0, b = (1.0 - s) / 4.0;
        c->nq = 4;
        for (int q = 0; q < 4; ++q) c->qw[q] = 1.0 / 24;
        const double pts[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 3; ++k) c->qp[q][k] = pts[q][k];
    }
    return CMX_OK;
}

int build_gate_table(cmx_ctx* c)
{
    if (c->precision != CMX_MIXED || c->integrator != CMX_RUSH_LARSEN || !c->cell_groups.empty() || c->split)
        return fail(c, CMX_E_ARG, "lookup tables need mixed precision, Rush-Larsen cells of one model and the ICI step");
    std::vector<float> table;
    const char* err = gate_table(c->model, c->dt_ode, c->cm, table, &c->lut_rows);
    if (err) return fail(c, CMX_E_ARG, err);
    c->lut_vmin = kLutVmin;
    c->lut_dv = kLutDv;
    return upload(c, &c->lut, table.data(), table.size());
}

int create_impl(const cmx_mesh* mesh, const cmx_tissue* tissue, const cmx_opts* opts, cmx_ctx* c)
{
    if (!mesh || !tissue || !opts) return fail(c, CMX_E_ARG, "null mesh, tissue or opts");
    // Cell types: one (model, integrator) for every node, or per-node groups.
    const bool grouped = tissue->n_groups > 1;
    std::vector<int32_t> group_of;
    if (grouped) {
        if (!tissue->group_of || !tissue->group_model || !tissue->group_integrator)
            return fail(c, CMX_E_ARG, "n_groups > 1 needs group_of, group_model and group_integrator");
        if (tissue->n_params > 0) return fail(c, CMX_E_ARG, "per-node parameters need a single cell type");
        if (opts->state_variable_interpolation)
            return fail(c, CMX_E_ARG, "state-variable interpolation needs a single cell type");
        std::vector<int> count(tissue->n_groups, 0);
        group_of.assign(tissue->group_of, tissue->group_of + std::max(mesh->n_nodes, 0));
        for (int32_t g : group_of) {
            if (g < 0 || g >= tissue->n_groups) return fail(c, CMX_E_ARG, "group_of out of range");
            ++count[g];
        }
        c->S = 0;
        for (int g = 0; g < tissue->n_groups; ++g) {
            CellGroup cg;
            cg.model = tissue->group_model[g];
            cg.integrator = tissue->group_integrator[g];
            if (model_sizes(cg.model, &cg.S, &cg.P)) return fail(c, CMX_E_ARG, "unknown cell model in group_model");
            if (cg.integrator < CMX_FORWARD_EULER || cg.integrator > CMX_GRL2)
                return fail(c, CMX_E_ARG, "unknown integrator in group_integrator");
            if (!model_supports(cg.model, cg.integrator))
                return fail(c, CMX_E_ARG, "this cell model has forward-Euler and Rush-Larsen forms only");
            cg.n = count[g];
            if (cg.n == 0) continue;
            c->S = std::max(c->S, cg.S);
            c->cell_groups.push_back(cg);
            c->cell_groups.back().h_node.reserve(cg.n);
        }
        // group indices of the non-empty groups, in order
        std::vector<int> slot(tissue->n_groups, -1);
        for (int g = 0, k = 0; g < tissue->n_groups; ++g)
            if (count[g]) slot[g] = k++;
        for (int i = 0; i < mesh->n_nodes; ++i) c->cell_groups[slot[group_of[i]]].h_node.push_back(i);
        if (c->cell_groups.size() == 1) {   // every node in one group: the plain layout
            const CellGroup only = c->cell_groups[0];
            c->cell_groups.clear();
            c->model = only.model;
            c->integrator = only.integrator;
        } else {
            c->model = c->cell_groups[0].model;
            c->integrator = c->cell_groups[0].integrator;
        }
        int S0;
        model_sizes(c->model, &S0, &c->P);
        if (c->cell_groups.empty()) c->S = S0;
    } else {
        if (model_sizes(tissue->model, &c->S, &c->P)) return fail(c, CMX_E_ARG, "unknown cell model");
        if (tissue->integrator < CMX_FORWARD_EULER || tissue->integrator > CMX_GRL2)
            return fail(c, CMX_E_ARG, "unknown integrator");
        if (!model_supports(tissue->model, tissue->integrator))
            return fail(c, CMX_E_ARG, "this cell model has forward-Euler and Rush-Larsen forms only");
        c->model = tissue->model;
        c->integrator = tissue->integrator;
    }
    if (opts->precision != CMX_FP64 && opts->precision != CMX_MIXED) return fail(c, CMX_E_ARG, "unknown precision");
    if (!(opts->dt_pde > 0) || !(opts->dt_ode > 0) || opts->dt_ode > opts->dt_pde * (1 + 1e-12))
        return fail(c, CMX_E_ARG, "need 0 < dt_ode <= dt_pde");
    const double ratio = opts->dt_pde / opts->dt_ode;
    const int n_sub = (int)std::floor(ratio + 0.5);
    if (std::fabs(ratio - n_sub) > 1e-9 * ratio)
        return fail(c, CMX_E_ARG, "dt_ode must divide dt_pde");
    if (!(opts->am > 0) || !(opts->cm > 0)) return fail(c, CMX_E_ARG, "Am and Cm must be positive");
    if (opts->ksp_max_its <= 0) return fail(c, CMX_E_ARG, "ksp_max_its must be positive");
    if (opts->ksp_rtol < 0 || opts->ksp_atol < 0) return fail(c, CMX_E_ARG, "negative KSP tolerance");
    if (tissue->n_params < 0 || tissue->n_params > c->P || (tissue->n_params > 0 && !tissue->params))
        return fail(c, CMX_E_ARG, "bad per-node parameter planes");
    if (tissue->n_stimuli < 0 || tissue->n_stimuli > 255 || (tissue->n_stimuli > 0 && !tissue->stimuli))
        return fail(c, CMX_E_ARG, "bad stimulus table (at most 255 rows)");
    c->n = mesh->n_nodes;
    c->dim = mesh->dim;
    if (mesh->elem_dim > 0 && mesh->elem_dim != mesh->dim && opts->state_variable_interpolation)
        return fail(c, CMX_E_ARG, "state-variable interpolation needs elements of the node dimension");
    c->precision = opts->precision;
    c->dt = opts->dt_pde;
    c->dt_ode = opts->dt_ode;
    c->n_sub = n_sub;
    c->am = opts->am;
    c->cm = opts->cm;
    c->rtol = opts->ksp_rtol;
    c->atol = opts->ksp_atol;
    c->max_its = opts->ksp_max_its;
    c->refine = opts->refine_sweeps;
    c->fold_refine = opts->separate_refinement == 0;
    c->use_graph = opts->use_graph;
    c->graph_profiler_off = (opts->use_graph && profiler_attached()) ? 1 : 0;
    c->split = opts->operator_splitting ? 1 : 0;
    c->svi = opts->state_variable_interpolation ? 1 : 0;
    {
        // TimeStepper(0, dt/2, dt_ode) step count (global/src/TimeStepper.cpp:43-121)
        const double end = 0.5 * opts->dt_pde, eps = end > 1.0 ? DBL_EPSILON * end : DBL_EPSILON;
        int k = 0;
        double t = 0.0;
        while (t < end) {
            double nx = (k + 1) * opts->dt_ode;
            if (end - nx <= eps) nx = end;
            t = nx;
            ++k;
        }
        c->n_half = k;
    }
    if (opts->extrapolate < 0 || opts->extrapolate > 2) return fail(c, CMX_E_ARG, "extrapolate must be 0, 1 or 2");
    c->extrapolate = opts->extrapolate;
    c->n_param_planes = tissue->n_params;
    c->budget = std::min(8, c->max_its);
    c->refine_budget = 1;

    HostCsr csr;
    std::string err;
    if (!assemble_p1(*mesh, opts->mass_lumping != 0, csr, err)) return fail(c, CMX_E_ARG, err);
    c->nnz = csr.row_ptr[c->n];
    c->width = csr.max_row;
    const int n = c->n;

    CMX_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    int fmt = opts->format;
    StencilTable st;
    OffsetEll oell;
    if (fmt == CMX_FMT_AUTO || fmt == CMX_FMT_STENCIL) {
        if (build_stencil(csr, kMaxPatterns, kMaxStencil, st)) fmt = CMX_FMT_STENCIL;
        else if (fmt == CMX_FMT_STENCIL)
            return fail(c, CMX_E_ARG, "mesh needs more than 64 row patterns or 16 stencil offsets: use the ELL format");
        else fmt = CMX_FMT_ELL_OFFSETS;
    }
    if (fmt == CMX_FMT_ELL_OFFSETS) {
        if (!build_offset_ell(csr, 255, oell)) {
            if (opts->format == CMX_FMT_ELL_OFFSETS)
                return fail(c, CMX_E_ARG, "mesh rows need more than 255 column-offset sets: use the ELL format");
            fmt = CMX_FMT_ELL;
        }
    }
    if (fmt != CMX_FMT_STENCIL && fmt != CMX_FMT_ELL && fmt != CMX_FMT_ELL_OFFSETS)
        return fail(c, CMX_E_ARG, "unknown matrix format");
    if (fmt != CMX_FMT_STENCIL && ell_size(n, csr.max_row) > (size_t)INT32_MAX)
        return fail(c, CMX_E_ARG, "mesh too large for the ELL format's 32-bit value index");
    c->format = fmt;
    int rc;
    const double cA = c->am * c->cm / c->dt;
    if (fmt == CMX_FMT_STENCIL) {
        // pad the width to the kernel's compile-time width (4, 8 or 16)
        const int W = st.width <= 4 ? 4 : (st.width <= 8 ? 8 : 16);
        const int NPt = st.n_patterns;
        c->width = W;
        c->n_patterns = NPt;
        c->h_off.assign(W, 0);
        std::copy(st.off.begin(), st.off.end(), c->h_off.begin());
        for (int32_t o : c->h_off) c->pad = std::max(c->pad, o < 0 ? -o : o);
        // whole 256-byte units, so that every vector starts on a cache-line
        // boundary (streamed stores then write whole lines)
        c->pad = (c->pad + 63) / 64 * 64;
        std::vector<double> m64((size_t)NPt * W, 0.0), k64((size_t)NPt * W, 0.0), a64((size_t)NPt * W, 0.0);
        std::vector<float> a32((size_t)NPt * W, 0.0f);
        std::vector<double> dinv64(NPt);
        std::vector<float> dinv32(NPt);
        for (int q = 0; q < NPt; ++q) {
            double diag = 0;
            for (int k = 0; k < st.width; ++k) {
                const size_t src = (size_t)q * st.width + k, dst = (size_t)q * W + k;
                m64[dst] = st.m[src];
                k64[dst] = st.k[src];
                a64[dst] = (1.0 / c->dt) * (c->am * c->cm * st.m[src]) + st.k[src];
                a32[dst] = (float)a64[dst];
                if (st.off[k] == 0) diag = a64[dst];
            }
            dinv64[q] = 1.0 / diag;
            dinv32[q] = (float)dinv64[q];
        }
        if ((rc = upload(c, &c->pid, st.pid.data(), n))) return rc;
        if ((rc = upload(c, &c->m64, m64.data(), m64.size()))) return rc;
        if ((rc = upload(c, &c->k64, k64.data(), k64.size()))) return rc;
        if ((rc = upload(c, &c->a64, a64.data(), a64.size()))) return rc;
        if ((rc = upload(c, &c->a32, a32.data(), a32.size()))) return rc;
        if ((rc = upload(c, &c->dinv64, dinv64.data(), NPt))) return rc;
        if ((rc = upload(c, &c->dinv32, dinv32.data(), NPt))) return rc;
    } else if (fmt == CMX_FMT_ELL_OFFSETS) {
        const int W = oell.width;
        const size_t E = oell.m.size();
        c->width = W;
        c->n_patterns = oell.n_sets;
        // guard zones as wide as the largest offset (whole 256-byte units)
        c->pad = (oell.max_abs + 63) / 64 * 64;
        std::vector<double> a64(E), dinv64(n);
        std::vector<float> a32(E), dinv32(n);
        for (size_t s = 0; s < E; ++s) {
            a64[s] = (1.0 / c->dt) * (c->am * c->cm * oell.m[s]) + oell.k[s];
            a32[s] = (float)a64[s];
        }
        for (int i = 0; i < n; ++i) {
            for (int p = csr.row_ptr[i]; p < csr.row_ptr[i + 1]; ++p)
                if (csr.col[p] == i) dinv64[i] = 1.0 / (cA * csr.mass[p] + csr.stiff[p]);
            dinv32[i] = (float)dinv64[i];
        }
        if ((rc = upload(c, &c->ell_set, oell.set.data(), n))) return rc;
        if ((rc = upload(c, &c->ell_off, oell.offsets.data(), oell.offsets.size()))) return rc;
        if ((rc = upload(c, &c->m64, oell.m.data(), E))) return rc;
        if ((rc = upload(c, &c->k64, oell.k.data(), E))) return rc;
        if ((rc = upload(c, &c->a64, a64.data(), E))) return rc;
        if ((rc = upload(c, &c->a32, a32.data(), E))) return rc;
        if ((rc = upload(c, &c->dinv64, dinv64.data(), n))) return rc;
        if ((rc = upload(c, &c->dinv32, dinv32.data(), n))) return rc;
    } else {
        EllMatrix ell;
        build_ell(csr, ell);
        const size_t E = ell.m.size();
        std::vector<double> a64(E), dinv64(n);
        std::vector<float> a32(E), dinv32(n);
        for (size_t s = 0; s < E; ++s) {
            a64[s] = (1.0 / c->dt) * (c->am * c->cm * ell.m[s]) + ell.k[s];
            a32[s] = (float)a64[s];
        }
        for (int i = 0; i < n; ++i) {
            for (int p = csr.row_ptr[i]; p < csr.row_ptr[i + 1]; ++p)
                if (csr.col[p] == i) dinv64[i] = 1.0 / (cA * csr.mass[p] + csr.stiff[p]);
            dinv32[i] = (float)dinv64[i];
        }
        if ((rc = upload(c, &c->ell_col, ell.col.data(), E))) return rc;
        if ((rc = upload(c, &c->m64, ell.m.data(), E))) return rc;
        if ((rc = upload(c, &c->k64, ell.k.data(), E))) return rc;
        if ((rc = upload(c, &c->a64, a64.data(), E))) return rc;
        if ((rc = upload(c, &c->a32, a32.data(), E))) return rc;
        if ((rc = upload(c, &c->dinv64, dinv64.data(), n))) return rc;
        if ((rc = upload(c, &c->dinv32, dinv32.data(), n))) return rc;
    }
    c->G = default_solver_blocks(c);

    // Cells: default initial state (AbstractCardiacCell::Init ->
    // ResetToInitialConditions), V from the model's initial voltage.
    const bool mixed = c->precision == CMX_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    const size_t rv = rs;
    const double* y0 = model_init(c->model);
    std::vector<double> V0(n, y0[0]);
    for (const CellGroup& g : c->cell_groups)
        for (int32_t i : g.h_node) V0[i] = model_init(g.model)[0];
    if ((rc = dalloc_vec(c, &c->V, n))) return rc;
    CMX_HIP(c, hipMemcpy(c->V, V0.data(), n * sizeof(double), hipMemcpyHostToDevice));
    c->ps = (n + 63) / 64 * 64;
    for (CellGroup& g : c->cell_groups) {
        g.ps = (g.n + 63) / 64 * 64;
        const double* gi = model_init(g.model);
        std::vector<double> gp((size_t)(g.S - 1) * g.n);
        for (int s = 1; s < g.S; ++s) std::fill(gp.begin() + (size_t)(s - 1) * g.n, gp.begin() + (size_t)s * g.n, gi[s]);
        if ((rc = dalloc(c, (char**)&g.state, (size_t)(g.S - 1) * g.ps * rs))) return rc;
        if ((rc = mixed ? to_device_planes<float>(c, g.state, gp.data(), g.S - 1, g.n, g.ps)
                        : to_device_planes<double>(c, g.state, gp.data(), g.S - 1, g.n, g.ps)))
            return rc;
        const int Pg = std::max(g.P, 1);
        std::vector<double> pd(model_pdef(g.model), model_pdef(g.model) + Pg);
        if ((rc = dalloc(c, (char**)&g.pdef, Pg * rs))) return rc;
        if ((rc = mixed ? to_device_planes<float>(c, g.pdef, pd.data(), 1, Pg, Pg)
                        : to_device_planes<double>(c, g.pdef, pd.data(), 1, Pg, Pg)))
            return rc;
        if ((rc = upload(c, &g.node, g.h_node.data(), g.h_node.size()))) return rc;
    }
    std::vector<double> planes(c->cell_groups.empty() ? (size_t)(c->S - 1) * n : 0);
    for (int s = 1; s < c->S && c->cell_groups.empty(); ++s)
        std::fill(planes.begin() + (size_t)(s - 1) * n, planes.begin() + (size_t)s * n, y0[s]);
    const size_t ns = c->cell_groups.empty() ? (size_t)(c->S - 1) : 0;
    if (ns) {
        if ((rc = dalloc(c, (char**)&c->state, ns * c->ps * rs))) return rc;
        if ((rc = mixed ? to_device_planes<float>(c, c->state, planes.data(), ns, n, c->ps)
                        : to_device_planes<double>(c, c->state, planes.data(), ns, n, c->ps)))
            return rc;
    }
    const int P = std::max(c->P, 1);
    std::vector<double> pdef(model_pdef(c->model), model_pdef(c->model) + P);
    if ((rc = dalloc(c, (char**)&c->pdef, P * rs))) return rc;
    if ((rc = mixed ? to_device_planes<float>(c, c->pdef, pdef.data(), 1, P, P)
                    : to_device_planes<double>(c, c->pdef, pdef.data(), 1, P, P)))
        return rc;
    if (c->n_param_planes > 0) {
        const size_t np = (size_t)c->n_param_planes;
        if ((rc = dalloc(c, (char**)&c->params, np * c->ps * rs))) return rc;
        c->param_planes_alloc = c->n_param_planes;
        if ((rc = mixed ? to_device_planes<float>(c, c->params, tissue->params, np, n, c->ps)
                        : to_device_planes<double>(c, c->params, tissue->params, np, n, c->ps)))
            return rc;
    }
    // Stimulus table (durations padded as the reference's constructors do).
    c->h_stim.resize(tissue->n_stimuli);
    for (int k = 0; k < tissue->n_stimuli; ++k) {
        cmx_stimulus s = tissue->stimuli[k];
        if (!pad_stimulus(s)) return fail(c, CMX_E_ARG, "stimulus duration exceeds its period");
        c->h_stim[k] = s;
    }
    c->h_stim_id.assign(n, 0);
    if (tissue->stim_index) {
        for (int i = 0; i < n; ++i) {
            const int k = tissue->stim_index[i];
            if (k < -1 || k >= tissue->n_stimuli) return fail(c, CMX_E_ARG, "stim_index out of range");
            c->h_stim_id[i] = (uint8_t)(k + 1);
        }
    }
    if ((rc = upload(c, &c->stim_id, c->h_stim_id.data(), n))) return rc;
    if ((rc = upload(c, &c->stim, c->h_stim.data(), c->h_stim.size()))) return rc;

    // Work vectors.
    auto rvec = [&](void** ptr) {
        return mixed ? dalloc_vec(c, (float**)ptr, n) : dalloc_vec(c, (double**)ptr, n);
    };
    if ((rc = rvec(&c->F))) return rc;
    for (int v = 0; v < V_COUNT; ++v)
        if ((rc = rvec(&c->vec[v]))) return rc;
    if (c->precision == CMX_MIXED && c->refine > 0) {
        if ((rc = dalloc(c, &c->r0_64, n))) return rc;
    }
    if (c->extrapolate) {
        if ((rc = dalloc_vec(c, &c->Y, n))) return rc;
        if ((rc = rvec(&c->G1))) return rc;
    }
    if (c->svi && !c->split && (rc = setup_svi(c, mesh, rvec))) return rc;
    if ((rc = dalloc(c, &c->partials, kPartBufs * (size_t)kMaxBlocks * 3))) return rc;
    CMX_HIP(c, hipMemset(c->partials, 0, kPartBufs * (size_t)kMaxBlocks * 3 * sizeof(double)));
    if ((rc = dalloc(c, &c->clock, 1))) return rc;
    if ((rc = dalloc(c, &c->ctl, 2))) return rc;
    if ((rc = dalloc(c, &c->hist, (size_t)c->max_its + 2))) return rc;
    if ((rc = dalloc(c, &c->its_log, 4 * (size_t)c->batch))) return rc;
    CMX_HIP(c, hipMemset(c->its_log, 0, 4 * c->batch * sizeof(int32_t)));
    SolveCtl z{};
    z.max_its = c->max_its;
    SolveCtl z2[2] = {z, z};
    CMX_HIP(c, hipMemcpy(c->ctl, z2, sizeof(z2), hipMemcpyHostToDevice));
    StepClock ck{};
    CMX_HIP(c, hipMemcpy(c->clock, &ck, sizeof(ck), hipMemcpyHostToDevice));
    CMX_HIP(c, hipHostMalloc((void**)&c->h_clock, sizeof(StepClock), hipHostMallocDefault));
    CMX_HIP(c, hipHostMalloc((void**)&c->h_ctl, 2 * sizeof(SolveCtl), hipHostMallocDefault));
    CMX_HIP(c, hipHostMalloc((void**)&c->h_its, 4 * c->batch * sizeof(int32_t), hipHostMallocDefault));
    *c->h_clock = ck;
    if (opts->lookup_tables) {
        rc = build_gate_table(c);
        if (rc) return rc;
    }
    CMX_HIP(c, hipDeviceSynchronize());
    return CMX_OK;
}

// Iion at the stored state of n_cells cells (cell j at node node[j], or j).
void launch_iionic(cmx_ctx* c, int model, int n_cells, const int32_t* node, int ps, const void* state,
                   const void* params, const void* pdef, int n_param_planes)
{
    const ModelOps* ops = model_ops(model);
    (c->precision == CMX_MIXED ? ops->iionic32 : ops->iionic64)(c->stream, grid_of(n_cells), n_cells, node, ps, c->V,
                                                                state, params, pdef, n_param_planes, c->cm, c->iion);
}

void release(cmx_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_times(c);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    if (c->h_clock) (void)hipHostFree(c->h_clock);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_its) (void)hipHostFree(c->h_its);
    c->h_clock = nullptr;
    c->h_ctl = nullptr;
    c->h_its = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

}  // namespace

extern "C" {

int cmx_create(const cmx_mesh* mesh, const cmx_tissue* tissue, const cmx_opts* opts, cmx_ctx** out)
{
    if (!out) return fail(nullptr, CMX_E_ARG, "null output pointer");
    *out = nullptr;
    cmx_ctx* c = new cmx_ctx();
    const int rc = create_impl(mesh, tissue, opts, c);
    if (rc) {
        g_create_error = c->err;
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return CMX_OK;
}

int cmx_destroy(cmx_ctx* ctx)
{
    release(ctx);
    delete ctx;
    return CMX_OK;
}

const char* cmx_last_error(const cmx_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int cmx_step(cmx_ctx* c, double t0, int32_t n_steps)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    if (n_steps < 0) return fail(c, CMX_E_ARG, "negative step count");
    if (n_steps == 0) return CMX_OK;
    return cmx::step_range(c, t0, n_steps, 0, n_steps);
}

}  // extern "C"

// Steps k0 .. k1-1 of the run cmx_step(c, t0, n_total) takes, with its step
// times: TimeStepper(t0, t0 + n_total dt, dt), t_k = t0 + k dt, the end
// snapped (global/src/TimeStepper.cpp:106-121), evaluated on the device.
int cmx::plain_cg_start(cmx_ctx* c, const PlainCg& s, const double* b)
{
    if (c->precision != CMX_FP64) return fail(c, CMX_E_ARG, "plain solves need an FP64 context");
    if (c->format == CMX_FMT_STENCIL) {
        if (c->width == 4) return plain_start_rows(c, stencil_rows<4>(c), s, b);
        if (c->width == 8) return plain_start_rows(c, stencil_rows<8>(c), s, b);
        return plain_start_rows(c, stencil_rows<16>(c), s, b);
    }
    return plain_start_rows(c, ell_rows(c), s, b);
}

int cmx::plain_cg_iters(cmx_ctx* c, const PlainCg& s, int from, int count)
{
    if (c->precision != CMX_FP64) return fail(c, CMX_E_ARG, "plain solves need an FP64 context");
    if (c->format == CMX_FMT_STENCIL) {
        if (c->width == 4) return plain_iters_rows(c, stencil_rows<4>(c), s, from, count);
        if (c->width == 8) return plain_iters_rows(c, stencil_rows<8>(c), s, from, count);
        return plain_iters_rows(c, stencil_rows<16>(c), s, from, count);
    }
    return plain_iters_rows(c, ell_rows(c), s, from, count);
}

int cmx::step_range(cmx_ctx* c, double t0, int n_total, int k0, int k1)
{
    StepClock ck = run_clock(t0, c->dt, n_total);
    ck.k = k0;
    ck.n_steps = k1;
    *c->h_clock = ck;
    CMX_HIP(c, hipMemcpyAsync(c->clock, c->h_clock, sizeof(StepClock), hipMemcpyHostToDevice, c->stream));
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    const int rc = (c->precision == CMX_MIXED) ? steps_typed<float, float>(c) : steps_typed<double, double>(c);
    if (rc) return rc;
    // with operator splitting F carries no ionic current: Iion is evaluated at the state
    c->stepped = !c->split;
    c->last_t_next = k1 == n_total ? ck.t_end : t0 + k1 * c->dt;
    return CMX_OK;
}

extern "C" {

int cmx_autotune(cmx_ctx* c, double t0, int32_t n_steps, int32_t n_candidates, const int32_t* candidates,
                 int32_t* chosen, double* ms_per_step)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    if (n_steps <= 0) return fail(c, CMX_E_ARG, "n_steps must be positive");
    const int cg_rows = (c->n + kCgBlock - 1) / kCgBlock;
    std::vector<int> cand;
    if (candidates && n_candidates > 0) {
        for (int k = 0; k < n_candidates; ++k) {
            if (candidates[k] < 1 || candidates[k] > kMaxBlocks)
                return fail(c, CMX_E_ARG, "candidate block counts must lie in [1, 1024]");
            cand.push_back(std::min(candidates[k], cg_rows));
        }
    } else {
        for (int g : {128, 192, 256, 320, 384, 512}) cand.push_back(std::min(g, cg_rows));
    }
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    // everything a trial run changes: the state, the initial-guess history,
    // activation times, budgets and statistics
    const size_t rv = c->precision == CMX_MIXED ? sizeof(float) : sizeof(double);
    const size_t rs = rv;
    std::vector<char> V(c->n * sizeof(double)), S(c->cell_groups.empty() ? (size_t)(c->S - 1) * c->ps * rs : 0),
        X0(c->n * rv), XP(c->n * rv),
        act(c->act ? 5 * (size_t)c->n * sizeof(double) : 0), Yv(c->Y ? c->n * sizeof(double) : 0), F(c->n * rv);
    auto copy_out = [&](std::vector<char>& h, const void* d) {
        return h.empty() ? hipSuccess : hipMemcpy(h.data(), d, h.size(), hipMemcpyDeviceToHost);
    };
    auto copy_in = [&](void* d, const std::vector<char>& h) {
        return h.empty() ? hipSuccess : hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice);
    };
    std::vector<std::vector<char>> GS;
    for (const CellGroup& g : c->cell_groups) GS.emplace_back((size_t)(g.S - 1) * g.ps * rs);
    CMX_HIP(c, copy_out(V, c->V));
    if (c->cell_groups.empty()) CMX_HIP(c, copy_out(S, c->state));
    for (size_t q = 0; q < GS.size(); ++q) CMX_HIP(c, copy_out(GS[q], c->cell_groups[q].state));
    CMX_HIP(c, copy_out(X0, c->vec[V_X0]));
    CMX_HIP(c, copy_out(XP, c->vec[V_XP]));
    CMX_HIP(c, copy_out(act, c->act));
    CMX_HIP(c, copy_out(Yv, c->Y));
    CMX_HIP(c, copy_out(F, c->F));   // the Iion cache is recovered from F
    const double last_t_next = c->last_t_next;
    const int budget = c->budget, refine_budget = c->refine_budget;
    const long long steps = c->steps, cg_total = c->cg_total, refine_total = c->refine_total, resumes = c->resumes,
                    folded_total = c->folded_total;
    const bool stepped = c->stepped;
    auto restore = [&]() -> int {
        CMX_HIP(c, copy_in(c->V, V));
        if (c->cell_groups.empty()) CMX_HIP(c, copy_in(c->state, S));
        for (size_t q = 0; q < GS.size(); ++q) CMX_HIP(c, copy_in(c->cell_groups[q].state, GS[q]));
        CMX_HIP(c, copy_in(c->vec[V_X0], X0));
        CMX_HIP(c, copy_in(c->vec[V_XP], XP));
        CMX_HIP(c, copy_in(c->act, act));
        CMX_HIP(c, copy_in(c->Y, Yv));
        CMX_HIP(c, copy_in(c->F, F));
        c->budget = budget;
        c->refine_budget = refine_budget;
        drop_graphs(c);
        return CMX_OK;
    };
    const int G0 = c->G, rows0 = c->cg_rows, gs0 = c->graph_steps;
    int best = G0, best_rows = rows0, best_gs = gs0, best_budget = budget, best_refine_budget = refine_budget;
    double best_ms = 0.0;
    int rc = CMX_OK;
    // one trial of the current launch configuration: an untimed run (graph
    // capture, caches), then the timed one, from the saved state each time
    auto trial = [&](double* out_ms) -> int {
        double ms = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            int r = restore();
            if (r) return r;
            CMX_HIP(c, hipDeviceSynchronize());
            const auto a = std::chrono::steady_clock::now();
            r = cmx::step_range(c, t0, n_steps, 0, n_steps);
            if (r) return r;
            CMX_HIP(c, hipStreamSynchronize(c->stream));
            ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count() / n_steps;
        }
        *out_ms = ms;
        return CMX_OK;
    };
    auto consider = [&](double ms, bool first) {
        if (first || ms < best_ms) {
            best_ms = ms;
            best = c->G;
            best_rows = c->cg_rows;
            best_gs = c->graph_steps;
            best_budget = c->budget;
            best_refine_budget = c->refine_budget;
        }
    };
    // coordinate search: solver blocks, then rows per thread in the CG
    // kernels (tile shape), then steps per captured graph (unroll depth)
    for (size_t k = 0; k < cand.size() && rc == CMX_OK; ++k) {
        c->G = cand[k];
        double ms = 0.0;
        if ((rc = trial(&ms))) break;
        if (ms_per_step) ms_per_step[k] = ms;
        consider(ms, k == 0);
    }
    for (int rows = 1; rows <= 2 && rc == CMX_OK; ++rows) {
        c->G = best;
        c->graph_steps = best_gs;
        c->cg_rows = rows;
        if (rows == (best_rows > 0 ? best_rows : (c->precision == CMX_MIXED ? 2 : 1))) continue;
        double ms = 0.0;
        if ((rc = trial(&ms))) break;
        consider(ms, false);
    }
    if (graphs_on(c))
        for (int gs = 4; gs <= kMaxGraphSteps && rc == CMX_OK; gs *= 2) {
            c->G = best;
            c->cg_rows = best_rows;
            c->graph_steps = gs;
            if (gs == best_gs) continue;
            double ms = 0.0;
            if ((rc = trial(&ms))) break;
            consider(ms, false);
        }
    c->G = rc == CMX_OK ? best : G0;
    c->cg_rows = rc == CMX_OK ? best_rows : rows0;
    c->graph_steps = rc == CMX_OK ? best_gs : gs0;
    int rc2 = restore();
    if (rc == CMX_OK) {
        // the budgets the trial of those same steps settled on
        c->budget = best_budget;
        c->refine_budget = best_refine_budget;
    }
    // the graphs of the next batch at the restored budgets, captured here
    // rather than inside the caller's next (possibly timed) cmx_step
    if (rc == CMX_OK && rc2 == CMX_OK)
        rc2 = c->precision == CMX_MIXED ? prime_typed<float, float>(c, c->batch) : prime_typed<double, double>(c, c->batch);
    c->steps = steps;
    c->cg_total = cg_total;
    c->refine_total = refine_total;
    c->resumes = resumes;
    c->folded_total = folded_total;
    c->stepped = stepped;
    c->last_t_next = last_t_next;
    if (rc) return rc;
    if (rc2) return rc2;
    if (chosen) *chosen = c->G;
    return CMX_OK;
}

int cmx_get_voltage(cmx_ctx* c, double* V)
{
    if (!c || !V) return fail(c, CMX_E_ARG, "null argument");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    CMX_HIP(c, hipMemcpy(V, c->V, c->n * sizeof(double), hipMemcpyDeviceToHost));
    return CMX_OK;
}

int cmx_set_voltage(cmx_ctx* c, const double* V)
{
    if (!c || !V) return fail(c, CMX_E_ARG, "null argument");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    CMX_HIP(c, hipMemcpy(c->V, V, c->n * sizeof(double), hipMemcpyHostToDevice));
    // the previous correction no longer predicts the next one
    const size_t rv = c->precision == CMX_MIXED ? sizeof(float) : sizeof(double);
    CMX_HIP(c, hipMemset(c->vec[V_X0], 0, c->n * rv));
    CMX_HIP(c, hipMemset(c->vec[V_XP], 0, c->n * rv));
    return CMX_OK;
}

int cmx_get_state(cmx_ctx* c, double* states)
{
    if (!c || !states) return fail(c, CMX_E_ARG, "null argument");
    int rc = cmx_get_voltage(c, states);
    if (rc) return rc;
    if (c->cell_groups.empty()) {
        const size_t ns = (size_t)(c->S - 1);
        return c->precision == CMX_MIXED ? from_device_planes<float>(c, states + c->n, c->state, ns, c->n, c->ps)
                                         : from_device_planes<double>(c, states + c->n, c->state, ns, c->n, c->ps);
    }
    // mixed cell types: node i's column holds its model's states, NaN past them
    const size_t n = (size_t)c->n;
    std::fill(states + n, states + (size_t)c->S * n, std::numeric_limits<double>::quiet_NaN());
    for (const CellGroup& g : c->cell_groups) {
        std::vector<double> planes((size_t)(g.S - 1) * g.n);
        rc = c->precision == CMX_MIXED ? from_device_planes<float>(c, planes.data(), g.state, g.S - 1, g.n, g.ps)
                                       : from_device_planes<double>(c, planes.data(), g.state, g.S - 1, g.n, g.ps);
        if (rc) return rc;
        for (int s = 1; s < g.S; ++s)
            for (int j = 0; j < g.n; ++j) states[(size_t)s * n + g.h_node[j]] = planes[(size_t)(s - 1) * g.n + j];
    }
    return CMX_OK;
}

int cmx_set_state(cmx_ctx* c, const double* states)
{
    if (!c || !states) return fail(c, CMX_E_ARG, "null argument");
    int rc = cmx_set_voltage(c, states);
    if (rc) return rc;
    c->stepped = false;
    if (c->cell_groups.empty()) {
        const size_t ns = (size_t)(c->S - 1);
        return c->precision == CMX_MIXED ? to_device_planes<float>(c, c->state, states + c->n, ns, c->n, c->ps)
                                         : to_device_planes<double>(c, c->state, states + c->n, ns, c->n, c->ps);
    }
    const size_t n = (size_t)c->n;
    for (const CellGroup& g : c->cell_groups) {
        std::vector<double> planes((size_t)(g.S - 1) * g.n);
        for (int s = 1; s < g.S; ++s)
            for (int j = 0; j < g.n; ++j) planes[(size_t)(s - 1) * g.n + j] = states[(size_t)s * n + g.h_node[j]];
        rc = c->precision == CMX_MIXED ? to_device_planes<float>(c, g.state, planes.data(), g.S - 1, g.n, g.ps)
                                       : to_device_planes<double>(c, g.state, planes.data(), g.S - 1, g.n, g.ps);
        if (rc) return rc;
    }
    return CMX_OK;
}

int cmx_get_iionic(cmx_ctx* c, double* out)
{
    if (!c || !out) return fail(c, CMX_E_ARG, "null argument");
    if (!c->stepped) {
        // No step since the state was set: evaluate at the current state.
        if (!c->iion) {
            int rc = dalloc(c, &c->iion, c->n);
            if (rc) return rc;
        }
        if (c->cell_groups.empty()) {
            launch_iionic(c, c->model, c->n, nullptr, c->ps, c->state, c->params, c->pdef, c->n_param_planes);
        } else {
            for (const CellGroup& g : c->cell_groups)
                launch_iionic(c, g.model, g.n, g.node, g.ps, g.state, nullptr, g.pdef, 0);
        }
        CMX_HIP(c, hipGetLastError());
        CMX_HIP(c, hipStreamSynchronize(c->stream));
        CMX_HIP(c, hipMemcpy(out, c->iion, c->n * sizeof(double), hipMemcpyDeviceToHost));
        return CMX_OK;
    }
    // After a step the cache is recovered from F = -Am Iion - Istim(t_next).
    std::vector<double> F(c->n);
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    int rc = c->precision == CMX_MIXED ? from_device_planes<float>(c, F.data(), c->F, 1, c->n, c->n)
                                       : from_device_planes<double>(c, F.data(), c->F, 1, c->n, c->n);
    if (rc) return rc;
    for (int i = 0; i < c->n; ++i) {
        const int sid = c->h_stim_id[i];
        const double is = sid ? host_stimulus(c->h_stim[sid - 1], c->last_t_next) : 0.0;
        out[i] = -(F[i] + is) / c->am;
    }
    return CMX_OK;
}

int cmx_set_activation(cmx_ctx* c, double threshold, int32_t every_steps, double t_now)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    if (every_steps <= 0) {
        c->act_every = 0;
        return CMX_OK;
    }
    if (!std::isfinite(threshold)) return fail(c, CMX_E_ARG, "activation threshold must be finite");
    if (!c->act) {
        int rc = dalloc(c, &c->act, 5 * (size_t)c->n);
        if (rc) return rc;
    }
    c->act_thr = threshold;
    c->act_every = every_steps;
    activation_init_kernel<<<grid_of(c->n), kBlock, 0, c->stream>>>(c->n, c->V, c->act, threshold, t_now);
    CMX_HIP(c, hipGetLastError());
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    return CMX_OK;
}

int cmx_get_activation(cmx_ctx* c, double* out)
{
    if (!c || !out) return fail(c, CMX_E_ARG, "null argument");
    if (!c->act) return fail(c, CMX_E_STATE, "activation tracking was never enabled (cmx_set_activation)");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    CMX_HIP(c, hipMemcpy(out, c->act, 5 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost));
    return CMX_OK;
}

int cmx_get_stats(cmx_ctx* c, cmx_stats* s)
{
    if (!c || !s) return fail(c, CMX_E_ARG, "null argument");
    s->steps = c->steps;
    s->cg_iterations = c->cg_total;
    s->refine_sweeps = c->refine_total;
    s->last_iterations = c->last_its;
    s->format = c->format;
    s->n_nodes = c->n;
    s->n_states = c->S;
    s->row_width = c->width;
    s->n_patterns = c->n_patterns;
    s->nnz = c->nnz;
    s->last_residual = c->last_residual;
    s->device_bytes = c->device_bytes;
    s->graph_launches = c->graph_launches;
    s->graph_captures = c->graph_captures;
    s->graph_mode = !c->use_graph ? 0 : (c->graph_profiler_off ? 2 : 1);
    s->solver_blocks = c->G;
    s->resumes = c->resumes;
    s->folded_refinements = c->folded_total;
    s->adjoint_slots = c->adj_slots;
    s->adjoint_sub_interval = c->adj_sub;
    s->cg_rows_per_thread = c->cg_rows > 0 ? c->cg_rows : (c->precision == CMX_MIXED ? 2 : 1);
    s->graph_steps = c->graph_steps;
    s->adjoint_iterations = c->adj_iterations;
    return CMX_OK;
}

int cmx_set_launch_config(cmx_ctx* c, int32_t solver_blocks, int32_t cg_rows, int32_t graph_steps)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    const int cg_max = (c->n + kCgBlock - 1) / kCgBlock;
    if (solver_blocks < 0 || solver_blocks > kMaxBlocks) return fail(c, CMX_E_ARG, "solver_blocks must lie in [0, 1024]");
    if (cg_rows < 0 || cg_rows > 2) return fail(c, CMX_E_ARG, "cg_rows must be 0, 1 or 2");
    if (graph_steps < 0 || graph_steps > kMaxGraphSteps || (graph_steps & (graph_steps - 1)))
        return fail(c, CMX_E_ARG, "graph_steps must be 0 or a power of two up to 32");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    if (solver_blocks > 0) c->G = std::min(solver_blocks, cg_max);
    if (cg_rows > 0) c->cg_rows = cg_rows;
    if (graph_steps > 0) c->graph_steps = graph_steps;
    return CMX_OK;
}

int cmx_set_profiling(cmx_ctx* c, int32_t mask)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    c->prof_mask = mask;
    return CMX_OK;
}

int cmx_get_kernel_times(cmx_ctx* c, double* total_ms, int64_t* launches, int32_t reset)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    collect_times(c);
    for (int k = 0; k < CMX_K_COUNT; ++k) {
        if (total_ms) total_ms[k] = c->k_ms[k];
        if (launches) launches[k] = c->k_launches[k];
        if (reset) { c->k_ms[k] = 0; c->k_launches[k] = 0; }
    }
    return CMX_OK;
}

int cmx_load_model(const char* path, int32_t* model_id)
{
    if (!path || !model_id) return fail(nullptr, CMX_E_ARG, "null argument");
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        const char* e = dlerror();
        return fail(nullptr, CMX_E_ARG, std::string("cannot load the cell-model plugin: ") + (e ? e : path));
    }
    using AbiFn = uint32_t (*)();
    using OpsFn = const ModelOps* (*)();
    const AbiFn abi = (AbiFn)dlsym(h, "cmx_model_plugin_abi");
    const OpsFn ops_of = (OpsFn)dlsym(h, "cmx_model_plugin_ops");
    if (!abi || !ops_of) {
        dlclose(h);
        return fail(nullptr, CMX_E_ARG, std::string(path) + " is not a cell-model plugin (no cmx_model_plugin_ops)");
    }
    if (abi() != kModelPluginAbi) {
        dlclose(h);
        return fail(nullptr, CMX_E_ARG, std::string(path) + " was built against other headers than this library");
    }
    const ModelOps* ops = ops_of();
    if (!ops || ops->meta.S < 2 || ops->meta.S > kMaxModelStates || ops->meta.P < 0 || !ops->cells64 || !ops->cells32) {
        dlclose(h);
        return fail(nullptr, CMX_E_ARG, std::string(path) + ": the model's table entry is incomplete or too large");
    }
    std::lock_guard<std::mutex> lock(g_plugin_mutex);
    if (g_plugin_count >= kMaxPluginModels) {
        dlclose(h);
        return fail(nullptr, CMX_E_ARG, "too many cell-model plugins loaded");
    }
    // the plugin stays loaded until the process ends: its kernels may sit in
    // captured graphs and in contexts that outlive this call
    const int k = g_plugin_count++;
    g_plugin_models[k].store(ops, std::memory_order_release);
    *model_id = kPluginModelBase + k;
    return CMX_OK;
}

int cmx_model_info(int32_t model, int32_t* n_states, int32_t* n_params)
{
    int S, P;
    if (model_sizes(model, &S, &P)) return fail(nullptr, CMX_E_ARG, "unknown cell model");
    if (n_states) *n_states = S;
    if (n_params) *n_params = P;
    return CMX_OK;
}

int cmx_gate_table(int32_t model, double dt, float* table, int32_t* rows, int32_t* width, double* vmin, double* dv)
{
    if (!(dt > 0)) return fail(nullptr, CMX_E_ARG, "dt must be positive");
    std::vector<float> t;
    int r = 0;
    const char* err = gate_table(model, dt, 1.0, t, &r);
    if (err) return fail(nullptr, CMX_E_ARG, err);
    if (rows) *rows = r;
    if (width) *width = (int32_t)(t.size() / (size_t)r);
    if (vmin) *vmin = kLutVmin;
    if (dv) *dv = kLutDv;
    if (table) std::copy(t.begin(), t.end(), table);
    return CMX_OK;
}

int cmx_model_param(int32_t model, int32_t index, const char** name, double* default_value)
{
    int S, P;
    if (model_sizes(model, &S, &P) || index < 0 || index >= P) return fail(nullptr, CMX_E_ARG, "bad model/parameter index");
    if (name) { ModelMeta m{}; model_meta(model, &m); *name = m.param_names[index]; }
    if (default_value) *default_value = model_pdef(model)[index];
    return CMX_OK;
}

int cmx_model_param_alias(int32_t model, int32_t index, const char** alias)
{
    int S, P;
    if (model_sizes(model, &S, &P) || index < 0 || index >= P || !alias)
        return fail(nullptr, CMX_E_ARG, "bad model/parameter index");
    ModelMeta m{};
    model_meta(model, &m);
    *alias = m.param_aliases ? m.param_aliases[index] : "";
    return CMX_OK;
}

int cmx_model_find_param(int32_t model, const char* name, int32_t* index)
{
    int S, P;
    if (model_sizes(model, &S, &P) || !name || !index) return fail(nullptr, CMX_E_ARG, "bad model or null argument");
    ModelMeta m{};
    model_meta(model, &m);
    for (int q = 0; q < P; ++q)
        if (!std::strcmp(name, m.param_names[q]) || (m.param_aliases && m.param_aliases[q][0] &&
                                                     !std::strcmp(name, m.param_aliases[q]))) {
            *index = q;
            return CMX_OK;
        }
    return fail(nullptr, CMX_E_ARG, std::string("No parameter named '") + name + "'");
}

int cmx_set_params(cmx_ctx* c, int32_t n_params, const double* params)
{
    if (!c) return fail(nullptr, CMX_E_ARG, "null context");
    if (n_params < 0 || n_params > c->P || (n_params > 0 && !params))
        return fail(c, CMX_E_ARG, "n_params must be in [0, the model's parameter count] with params given");
    if (!c->cell_groups.empty()) return fail(c, CMX_E_ARG, "per-node parameters need a single cell type");
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    const bool mixed = c->precision == CMX_MIXED;
    if (n_params > c->param_planes_alloc) {
        int rc = dalloc(c, (char**)&c->params, (size_t)n_params * c->ps * (mixed ? sizeof(float) : sizeof(double)));
        if (rc) return rc;
        c->param_planes_alloc = n_params;
    }
    if (n_params > 0) {
        int rc = mixed ? to_device_planes<float>(c, c->params, params, n_params, c->n, c->ps)
                       : to_device_planes<double>(c, c->params, params, n_params, c->n, c->ps);
        if (rc) return rc;
    }
    c->n_param_planes = n_params;
    // captured steps hold the parameter planes' address and count
    drop_graphs(c);
    return CMX_OK;
}

int cmx_model_state_name(int32_t model, int32_t index, const char** name)
{
    int S, P;
    if (model_sizes(model, &S, &P) || index < 0 || index >= S || !name) return fail(nullptr, CMX_E_ARG, "bad model/index");
    ModelMeta m{};
    model_meta(model, &m);
    *name = m.state_names[index];
    return CMX_OK;
}

int cmx_model_initial_state(int32_t model, double* y)
{
    int S, P;
    if (model_sizes(model, &S, &P) || !y) return fail(nullptr, CMX_E_ARG, "bad model");
    std::copy(model_init(model), model_init(model) + S, y);
    return CMX_OK;
}

int cmx_slab_mesh(int32_t dim, int32_t ex, int32_t ey, int32_t ez, double h,
                  int32_t* n_nodes, int32_t* n_elems, double* coords, int32_t* elems)
{
    const int rc = slab_mesh(dim, ex, ey, ez, h, n_nodes, n_elems, coords, elems);
    if (rc) return fail(nullptr, rc, "bad slab dimensions");
    return CMX_OK;
}

}  // extern "C"
