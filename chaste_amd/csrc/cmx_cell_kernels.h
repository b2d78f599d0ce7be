// Cell-model kernels of the tissue step: the cell kernel in its roles (ICI
// step fused with the forcing and the settling of the previous step,
// operator-splitting half steps), the Iion evaluation, the state-variable
// interpolation correction, and their typed launches.  They are templates on
// the model; each model's translation unit (cmx_model_*.hip) instantiates
// them and fills its ModelOps entry, so the runtime reaches every model
// through one table.
#pragma once
#include <type_traits>

#include "cmx_launch.h"

namespace cmx {

// What settles a pending step (its final solve decided, V^{n+1} written):
// the cell kernel of the next step, or the finish kernel at a batch's end.
template <class RV>
struct SettleArgs {
    SolveCtl* ctl;                     // final solve of the pending step
    SolveCtl* ctl0;                    // first solve, when the reduction's slot 1 tells whether its
                                       // recursive residual converged (refinement folded in); or null
    const double* __restrict__ pin;    // the last reduction of that solve
    int gin;                           // its slot count
    const double* __restrict__ Z;      // V^n, or V^n + x0 with an initial guess
    const RV* __restrict__ x1;         // first solve's correction
    const RV* __restrict__ x2;         // refinement solve's correction, or null
    RV* __restrict__ guess;            // x0 of the pending step, replaced by the next step's, or null
    RV* __restrict__ prev;             // second order: x_{n-2}, replaced by x_{n-1}; or null
    double* __restrict__ act;          // [5][n] activation times, or null
    double act_thr;
    int act_every;
};

// ActivationOutputModifier::ProcessSolutionAtTimeStep for node i at time t
// (heart/src/problem/ActivationOutputModifier.cpp:94-125).
__device__ __forceinline__ void modifier_update(double* __restrict__ act, int n, int i, double t, double v, double thr)
{
    double* a0 = act + i;
    const double fa = a0[0], fr = a0[n], sa = a0[2 * (size_t)n], sr = a0[3 * (size_t)n];
    if (v > thr && fa < 0.0) a0[0] = t;
    else if (fa >= 0.0 && fr < 0.0 && v < thr) a0[n] = t;
    else if (fr >= 0.0 && sa < 0.0 && v > thr) a0[2 * (size_t)n] = t;
    else if (sa >= 0.0 && sr < 0.0 && v < thr) a0[3 * (size_t)n] = t;
}

// First upward crossing of the threshold, linear in time between (t0, v0)
// and (t1, v1).
template <class RV>
__device__ __forceinline__ void record_crossing(const SettleArgs<RV>& s, int n, int i, double v0, double v1,
                                                double t0, double t1)
{
    double* cross = s.act + 4 * (size_t)n + i;
    if (*cross < 0.0 && v0 < s.act_thr && v1 >= s.act_thr) *cross = t0 + (t1 - t0) * (s.act_thr - v0) / (v1 - v0);
}

// Activation bookkeeping of a settled step k: v0 at time t0 (V^k at t_k, or
// with operator splitting the voltage after the first half step, at the
// midpoint) -> V^{k+1} = v1 at t_{k+1}.
template <class RV>
__device__ __forceinline__ void track_activation(const SettleArgs<RV>& s, int n, int i, const StepClock& clk, int k,
                                                 double v0, double v1, double t0)
{
    const double t1 = step_next_time(clk, k);
    record_crossing(s, n, i, v0, v1, t0, t1);
    if ((k + 1) % s.act_every == 0) modifier_update(s.act, n, i, t1, v1, s.act_thr);
}

// The next step's initial-guess correction from the settled step's total
// correction xt = x0 + x1 (+ x2): xt itself, or 2 xt - x_{n-2} (linear
// extrapolation of the corrections); updates the stored x_{n-2}.
template <class RV>
__device__ __forceinline__ RV next_guess(const SettleArgs<RV>& s, int i, RV xt, RV prev)
{
    if (!s.prev) return xt;
    s.prev[i] = xt;
    return RV(2) * xt - prev;
}

// The final solve converged earlier, or its last reduction passes the test
// (identical in every block).  Lead records a convergence found here.
__device__ __forceinline__ bool solve_converged(SolveCtl* ctl, int done, double tol2, double zeta, bool lead)
{
    if (done == 1) return true;
    if (done == 0 && zeta <= tol2) {
        if (lead) { ctl->done = 1; ctl->zz = zeta; }
        return true;
    }
    return false;
}

// Iteration counts of step k for the host: first solve, whether it
// converged (a refined step's first solve may hand over unconverged),
// refinement solve.
__device__ __forceinline__ void log_step(int32_t* its_log, int log_len, int k, const SolveCtl* ctl, bool refined)
{
    int32_t* slot = its_log + 4 * (k % log_len);
    slot[0] = ctl[0].it;
    slot[1] = refined ? (ctl[0].done == 1 ? 1 : 0) : 1;
    slot[2] = refined ? ctl[1].it : 0;
}

template <class RS, class RV>
struct CellArgs {
    int n;                     // nodes (V, F, activation planes)
    int n_cells;               // cells of this launch: thread j steps node node[j] (node == null: node j)
    const int32_t* __restrict__ node;
    int n_sub;                 // ODE substeps per PDE step
    double dt_ode;             // ms
    double am, cm;
    StepClock* clock;
    double* __restrict__ V;
    int ps;                    // plane stride (n rounded up to whole 256-byte units)
    RS* __restrict__ state;    // [NS-1][ps]: plane s-1 holds state s (V excluded), indexed by cell j
    const RS* __restrict__ params;  // [NP][ps] or null, indexed by cell j
    const RS* __restrict__ pdef;    // [NP] defaults
    int n_param_planes;
    const uint8_t* __restrict__ stim_id;  // 0 = none, k = row k-1
    const cmx_stimulus* __restrict__ stim;
    RV* __restrict__ F;        // out: -Am Iion - Istim(t_next)
    RV* __restrict__ iion_out; // out (state-variable interpolation): Iion, or null
    SettleArgs<RV> s;          // the previous step, when pending
    int n_half;                // operator splitting: ODE substeps per half step
    // CG initial guess V^n + x0 (x0 = the previous step's correction), when s.guess is set:
    double c;                       // Am Cm / dt
    double* __restrict__ Y;         // out: V^n + x0
    RV* __restrict__ G1;            // out: F - c x0
    // ICI steps k in [trace_k0, trace_k0 + trace_len): V^k and the states y^k
    // the step starts from, to trace + (k - trace_k0) * trace_slot as V at
    // [0, ps) and state s at [s*ps, (s+1)*ps) (the adjoint's replay), or null
    double* __restrict__ trace;
    int trace_k0, trace_len;
    size_t trace_slot;
    // gate lookup table (GateCols layout), rows at lut_vmin + r / lut_inv_dv
    const float* __restrict__ lut;
    float lut_vmin, lut_inv_dv;
    int lut_rows;
};

// Cell-kernel roles: the ICI step (V clamped over the step), the operator-
// splitting step (second half of the previous step and first half of this
// one, V updated), and the splitting step's last half only (end of a batch).
enum CellMode { kModeIci = 0, kModeSplit = 1, kModeSplitFinal = 2 };

// One backward-Euler substep of every state but V at voltage V
// (ComputeOneStepExceptVoltage of the reference's *BackwardEuler cells,
// heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:180-200), computed in
// FP64 whatever the state precision: Newton's update test (max-norm <= 1e-6)
// is below the FP32 spacing of the larger concentrations.  Returns false
// when Newton did not converge.
template <class Model, class RS>
__device__ __forceinline__ bool be_substep(RS* y, const RS* p, double V, double st, double cm, double h)
{
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    double yd[NS], pd[NP];
#pragma unroll
    for (int s = 0; s < NS; ++s) yd[s] = (double)y[s];
#pragma unroll
    for (int q = 0; q < NP; ++q) pd[q] = (double)p[q];
    yd[0] = V;
    const int its = Model::template be_step<double>(yd, V, st, pd, cm, h);
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = (RS)yd[s];
    return its >= 0;
}

// One generalized Rush-Larsen substep in FP64 whatever the state precision:
// the forward difference of step 1e-8 that gives the diagonal Jacobian
// entries vanishes below the FP32 spacing of most states.  V is y[0] on entry
// (FP64 vd); with update_v it is advanced into vd.
template <class Model, class RS, int ORDER>
__device__ __forceinline__ void grl_substep(RS* y, const RS* p, double& vd, double st, double cm, double h,
                                            bool update_v)
{
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    double yd[NS], pd[NP];
#pragma unroll
    for (int s = 1; s < NS; ++s) yd[s] = (double)y[s];
#pragma unroll
    for (int q = 0; q < NP; ++q) pd[q] = (double)p[q];
    yd[0] = vd;
    Model::template grl_step<double, ORDER>(yd, st, pd, cm, h, update_v);
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = (RS)yd[s];
    if (update_v) vd = yd[0];
}

// ODE substeps over [t0, t1] with the voltage updated too, as
// SolveAndUpdateState does (heart/src/odes/AbstractCardiacCell.cpp:70-73,
// ode/src/solver/AbstractOneStepIvpOdeSolver.cpp:93-153 for forward Euler,
// heart/src/odes/AbstractRushLarsenCardiacCell.cpp:130-153 for Rush-Larsen:
// every rate evaluated at the substep's start, V by forward Euler;
// heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:300-330 for backward
// Euler: V by forward Euler first, the other states then implicit at the new
// V; Rush-Larsen and backward-Euler cells always step by dt_ode).  vd is the
// FP64 voltage.
template <class Model, class RS, class RV, int IG>
__device__ __forceinline__ bool integrate_with_voltage(const CellArgs<RS, RV>& a, RS* y, const RS* p, int sid,
                                                       double& vd, double t0, double t1)
{
    constexpr int NS = Model::NS;
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    const RS cm = (RS)a.cm;
    bool ok = true;
    for (int ks = 0; ks < a.n_half; ++ks) {
        const double ts = t0 + ks * a.dt_ode;
        const double h = (IG == CMX_FORWARD_EULER && ks + 1 == a.n_half) ? t1 - ts : a.dt_ode;
        const double std_ = sid ? stimulus_at(a.stim[sid - 1], ts) / a.am : 0.0;
        const RS st = (RS)std_;
        const RS V = (RS)vd;
        y[0] = V;
        if constexpr (IG == CMX_BACKWARD_EULER) {
            vd += h * (double)Model::template dvdt<RS>(y, V, st, p, cm);
            ok &= be_substep<Model, RS>(y, p, vd, std_, a.cm, h);
            continue;
        }
        if constexpr (IG == CMX_GRL1 || IG == CMX_GRL2) {
            grl_substep<Model, RS, IG == CMX_GRL1 ? 1 : 2>(y, p, vd, std_, a.cm, h, true);
            continue;
        }
        RS dy[NS], gi[NS], gk[NS];
        Model::template eval<RS, RL>(y, V, st, p, cm, dy, gi, gk);
        const RS dv = Model::template dvdt<RS>(y, V, st, p, cm);
        const RS hh = (RS)h;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            if (RL && Model::gate(s)) y[s] = gi[s] + (y[s] - gi[s]) * cmx_decay(gk[s], hh);
            else y[s] += hh * dy[s];
        }
        vd += h * (double)dv;
    }
    return ok;
}

// One thread per node; the state lives in registers for all substeps.  When
// the previous step is pending, the kernel first settles it: the verdict on
// its final solve (a failed one poisons the clock and nothing is written),
// V = Z + x1 (+ x2) and the next initial guess; with operator splitting
// (OperatorSplittingMonodomainSolver, heart/src/solver/electrics/monodomain/
// OperatorSplittingMonodomainSolver.cpp:40-137) that voltage then runs the
// previous step's second half-step of the cells.
template <class Model, class RS, class RV, int IG, int MODE, bool GROUPED, bool LUT = false>
__global__ __launch_bounds__(kBlock) void cell_step_kernel(CellArgs<RS, RV> a)
{
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    static_assert(!LUT || (RL && MODE == kModeIci), "lookup tables serve Rush-Larsen ICI steps");
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    // Thread j steps cell j of the launch, which sits at node i (GROUPED:
    // node[j]; one cell group: i = j).  The prologue is branch-free (indices
    // clamped for the lanes past the last cell, absent vectors read through
    // a present one and masked) and issues its loads in the order they are
    // needed -- the reduction, the node-indexed settle vectors, the state
    // planes, the parameters -- so each wait covers only what it needs.
    const int j = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool live = j < a.n_cells;
    const int jl = live ? j : a.n_cells - 1;
    PartialLoad<3, kBlock> pl(a.s.pin, a.s.gin);
    const StepClock clk = *a.clock;
    const int done = a.s.ctl->done;
    const double tol2 = a.s.ctl->tol2;
    const int done0 = (a.s.ctl0 ? a.s.ctl0 : a.s.ctl)->done;
    const int i = GROUPED ? a.node[jl] : jl;
    const uint32_t ui = (uint32_t)i;
    const bool has_x2 = a.s.x2 != nullptr, has_guess = a.s.guess != nullptr, has_prev = a.s.prev != nullptr;
    const double z = ldu(a.s.Z, ui);
    const RV d1 = ldu(a.s.x1, ui);
    const RV d2 = ldu(has_x2 ? a.s.x2 : a.s.x1, ui);
    const RV g0 = ldu(has_guess ? (const RV*)a.s.guess : a.s.x1, ui);
    const RV gp = ldu(has_prev ? (const RV*)a.s.prev : a.s.x1, ui);
    const int sid = (int)ldu(a.stim_id, ui);
    RS y[NS];
    RS p[NP];
    // state planes through one buffer descriptor: plane s at scalar offset
    // (s-1) ps, every plane at the thread's one vector offset
    const __amdgpu_buffer_rsrc_t st_rs = raw_rsrc(a.state, (uint32_t)((NS - 1) * a.ps * sizeof(RS)));
    const uint32_t vj = (uint32_t)jl * (uint32_t)sizeof(RS);
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = buf_load<RS>(st_rs, vj, (uint32_t)((s - 1) * a.ps * sizeof(RS)));
#pragma unroll
    for (int q = 0; q < NP; ++q)
        p[q] = (q < a.n_param_planes) ? ldu(a.params + (size_t)q * a.ps, (uint32_t)jl)
                                      : (Model::NP > 0 ? a.pdef[q] : RS(0));
    const RV dx = d1 + (has_x2 ? d2 : RV(0));
    RV x0 = has_guess ? g0 : RV(0);
    const RV xp = has_prev ? gp : RV(0);
    double g[3];
    pl.total(g);
    if (clk.poison) return;
    if (MODE == kModeSplitFinal && !clk.pending) return;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    int k = clk.k;
    double vn = 0;
    if (clk.pending) {
        if (!solve_converged(a.s.ctl, done, tol2, g[2], lead)) {
            if (lead) {
                a.clock->poison = 1;
                a.clock->failed_step = clk.k;
            }
            return;
        }
        if (lead && a.s.ctl0 && done0 == 0 && g[1] <= tol2) a.s.ctl0->done = 1;
        vn = z + (double)dx;
        double v0 = z - (double)x0, t0 = step_time(clk, clk.k);
        if constexpr (MODE != kModeIci) {
            // second half-step of the pending step's cells, V updated
            if (live) {
                const double tp = step_time(clk, clk.k), tq = step_next_time(clk, clk.k);
                const double th = tp + 0.5 * (tq - tp);
                t0 = th;
                if (!integrate_with_voltage<Model, RS, RV, IG>(a, y, p, sid, vn, th, tq)) a.clock->nonfinite = 2;
            }
        }
        if (live && a.s.act) track_activation(a.s, a.n, i, clk, clk.k, v0, vn, t0);
        if (live && a.s.guess) x0 = next_guess(a.s, i, x0 + dx, xp);
        k = clk.k + 1;
    } else if (live) {
        vn = ldu(a.V, ui);
    }
    if (!live) return;
    if (clk.pending) {
        if (a.s.guess) stu(a.s.guess, ui, x0);
        if (!isfinite(vn)) a.clock->nonfinite = 1;
    }
    if constexpr (MODE == kModeSplitFinal) {
        stu(a.V, ui, vn);
#pragma unroll
        for (int s = 1; s < NS; ++s) buf_store<RS>(st_rs, vj, (uint32_t)((s - 1) * a.ps * sizeof(RS)), y[s]);
        return;
    }
    const double t = step_time(clk, k), t_next = step_next_time(clk, k);
    RV f;
    if constexpr (MODE == kModeSplit) {
        // first half-step of this step's cells, V updated; the diffusion
        // solve that follows has no reaction term
        const double v_start = vn, th = t + 0.5 * (t_next - t);
        if (!integrate_with_voltage<Model, RS, RV, IG>(a, y, p, sid, vn, t, th)) a.clock->nonfinite = 2;
        if (a.s.act) record_crossing(a.s, a.n, i, v_start, vn, t, th);
        stu(a.V, ui, vn);
        f = RV(0);
    } else {
        if (clk.pending) stu(a.V, ui, vn);
        if (a.trace && k >= a.trace_k0 && k < a.trace_k0 + a.trace_len) {
            double* tr = a.trace + (size_t)(k - a.trace_k0) * a.trace_slot;
            tr[i] = vn;
#pragma unroll
            for (int s = 1; s < NS; ++s) tr[(size_t)s * a.ps + i] = (double)y[s];
        }
        const RS V = (RS)vn;
        y[0] = V;
        const RS cm = (RS)a.cm;
        // TimeStepper(t, t_next, dt_ode): substep k starts at t + k*dt_ode
        // (global/src/TimeStepper.cpp:106-121); forward Euler takes the remaining
        // interval on the last substep, Rush-Larsen and backward Euler always
        // dt_ode (heart/src/odes/AbstractRushLarsenCardiacCell.cpp:106-128,
        // heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:204-230).
        bool ok = true;
        // V is clamped over the substeps: one table lookup serves them all
        // (the gate outputs of the model evaluation below are then unused)
        float tgi[NS], tgd[NS];
        if constexpr (LUT) {
            if (!lut_gates<Model>(a.lut, a.lut_vmin, a.lut_inv_dv, a.lut_rows, (float)V, tgi, tgd))
                a.clock->nonfinite = 4;
        }
        for (int ks = 0; ks < a.n_sub; ++ks) {
            const double ts = t + ks * a.dt_ode;
            const double h = (IG == CMX_FORWARD_EULER && ks + 1 == a.n_sub) ? t_next - ts : a.dt_ode;
            const double std_ = sid ? stimulus_at(a.stim[sid - 1], ts) / a.am : 0.0;
            const RS st = (RS)std_;
            if constexpr (IG == CMX_BACKWARD_EULER) {
                ok &= be_substep<Model, RS>(y, p, (double)V, std_, a.cm, h);
                continue;
            }
            if constexpr (IG == CMX_GRL1 || IG == CMX_GRL2) {
                double vclamp = vn;
                grl_substep<Model, RS, IG == CMX_GRL1 ? 1 : 2>(y, p, vclamp, std_, a.cm, h, false);
                continue;
            }
            RS dy[NS], gi[NS], gk[NS];
            Model::template eval<RS, RL>(y, V, st, p, cm, dy, gi, gk);
            const RS hh = (RS)h;
#pragma unroll
            for (int s = 1; s < NS; ++s) {
                if (LUT && Model::gate(s)) y[s] = tgi[s] + (y[s] - tgi[s]) * tgd[s];
                else if (RL && Model::gate(s)) y[s] = gi[s] + (y[s] - gi[s]) * cmx_decay(gk[s], hh);
                else y[s] += hh * dy[s];
            }
        }
        if (!ok) a.clock->nonfinite = 2;
        const RS iion = Model::template iion<RS>(y, V, p, cm);
        const double istim = sid ? stimulus_at(a.stim[sid - 1], t_next) : 0.0;
        f = (RV)(-a.am * (double)iion - istim);
        if (a.iion_out) stu(a.iion_out, ui, (RV)iion);
    }
    stu(a.F, ui, f);
    if (a.s.guess) {
        stu(a.Y, ui, vn + (double)x0);
        stu(a.G1, ui, (RV)((double)f - a.c * (double)x0));
    }
#pragma unroll
    for (int s = 1; s < NS; ++s) buf_store<RS>(st_rs, vj, (uint32_t)((s - 1) * a.ps * sizeof(RS)), y[s]);
}

// Iion at the stored state (cmx_get_iionic before any step); cell j of the
// group sits at node node[j] (node == null: node j).
template <class Model, class RS>
__global__ __launch_bounds__(kBlock) void iionic_kernel(int n_cells, const int32_t* __restrict__ node, int ps,
                                                        const double* __restrict__ V, const RS* __restrict__ state,
                                                        const RS* __restrict__ params, const RS* __restrict__ pdef,
                                                        int n_param_planes, double cm, double* __restrict__ out)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n_cells) return;
    const int i = node ? node[j] : j;
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    RS y[NS];
    y[0] = (RS)V[i];
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = state[(size_t)(s - 1) * ps + j];
    RS p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        p[q] = (q < n_param_planes) ? params[(size_t)q * ps + j] : (Model::NP > 0 ? pdef[q] : RS(0));
    out[i] = (double)Model::template iion<RS>(y, y[0], p, (RS)cm);
}

// State-variable interpolation (AbstractCorrectionTermAssembler.cpp:98-156,
// MonodomainCorrectionTermAssembler.cpp:48-74).  Every element whose nodal
// ionic currents differ pairwise by more than 1 uA/cm^2 adds, per quadrature
// point, w |J| phi_a (-Am) (Iion(sum_j phi_j y_j) - sum_j phi_j Iion_j) to the
// right-hand side of its node a (the interpolated state includes V; the cell
// parameters are those of the element's first node).
template <class RS, class RV>
struct SviArgs {
    int n, nb, nq;
    double am, cm;
    const StepClock* clock;
    const int32_t* __restrict__ elems;     // [ne][nb]
    const double* __restrict__ edet;       // [ne] |det J|
    const int32_t* __restrict__ inc_ptr;   // [n+1] node -> incident (element, local) list
    const int32_t* __restrict__ inc;       //   entries 4 e + local, ascending e
    const RV* __restrict__ iion;           // nodal Iion after the cell step
    const double* __restrict__ V;          // V^n
    const RS* __restrict__ state;          // [NS-1][ps]
    int ps;
    const RS* __restrict__ params;
    const RS* __restrict__ pdef;
    int n_param_planes;
    double qw[4];
    double qp[4][3];
    RV* __restrict__ out;                  // correction added to b
};

// Node a's share of element e's correction (0 when the element is below the
// criterion), in the reference's accumulation order.
template <class Model, class RS, class RV>
__device__ __forceinline__ double svi_element_term(const SviArgs<RS, RV>& a, int e, int loc)
{
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    int nd[4];
    double ii[4];
    for (int j = 0; j < a.nb; ++j) {
        nd[j] = a.elems[(size_t)e * a.nb + j];
        ii[j] = (double)a.iion[nd[j]];
    }
    double di = 0;
    for (int j = 0; j < a.nb; ++j)
        for (int k = j + 1; k < a.nb; ++k) di = fmax(di, fabs(ii[j] - ii[k]));
    if (!(di > 1.0)) return 0.0;
    const double det = a.edet[e];
    RS p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        p[q] = (q < a.n_param_planes) ? a.params[(size_t)q * a.ps + nd[0]] : (Model::NP > 0 ? a.pdef[q] : RS(0));
    double be = 0.0;
    for (int q = 0; q < a.nq; ++q) {
        double phi[4], sum = 0;
        for (int c = 0; c < a.nb - 1; ++c) {
            phi[c + 1] = a.qp[q][c];
            sum += a.qp[q][c];
        }
        phi[0] = 1.0 - sum;
        double yq[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) yq[s] = 0.0;
        double iq = 0.0;
        for (int j = 0; j < a.nb; ++j) {
            iq += phi[j] * ii[j];
            yq[0] += phi[j] * a.V[nd[j]];
#pragma unroll
            for (int s = 1; s < NS; ++s) yq[s] += phi[j] * (double)a.state[(size_t)(s - 1) * a.ps + nd[j]];
        }
        RS yr[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) yr[s] = (RS)yq[s];
        const double isv = (double)Model::template iion<RS>(yr, yr[0], p, (RS)a.cm);
        const double wJ = a.qw[q] * det;
        be += phi[loc] * (-a.am) * (isv - iq) * wJ;
    }
    return be;
}

// One thread per node.  A node takes part only if some element through it
// passes the criterion, so the range of Iion over its matrix row (every
// node it shares an element with) is checked first.
template <class Model, class RS, class RV, class Rows>
__global__ __launch_bounds__(kBlock) void svi_kernel(Rows R, SviArgs<RS, RV> a)
{
    const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= a.n || a.clock->poison) return;
    const auto row = R.row(i);
    double lo = (double)a.iion[i], hi = lo;
    for (int k = 0; k < R.width(); ++k) {
        const double v = (double)a.iion[R.col(row, k)];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    double acc = 0.0;
    if (hi - lo > 1.0) {
        const int q1 = a.inc_ptr[i + 1];
        for (int q = a.inc_ptr[i]; q < q1; ++q) {
            const int code = a.inc[q];
            acc += svi_element_term<Model, RS, RV>(a, code >> 2, code & 3);
        }
    }
    a.out[i] = (RV)acc;
}

// ----- typed launches -------------------------------------------------------
template <class Model, class RS, class RV, int IG, bool GROUPED>
void launch_cell_grouping(cmx_ctx* c, int mode, int kind, int G, const CellArgs<RS, RV>& a)
{
    if constexpr (IG == CMX_RUSH_LARSEN && std::is_same<RS, float>::value) {
        if (mode == kModeIci && a.lut) {
            CMX_LAUNCH(c, kind, (cell_step_kernel<Model, RS, RV, IG, kModeIci, GROUPED, true>), G, kBlock, a);
            return;
        }
    }
    if (mode == kModeIci) CMX_LAUNCH(c, kind, (cell_step_kernel<Model, RS, RV, IG, kModeIci, GROUPED>), G, kBlock, a);
    else if (mode == kModeSplit) CMX_LAUNCH(c, kind, (cell_step_kernel<Model, RS, RV, IG, kModeSplit, GROUPED>), G, kBlock, a);
    else CMX_LAUNCH(c, kind, (cell_step_kernel<Model, RS, RV, IG, kModeSplitFinal, GROUPED>), G, kBlock, a);
}

template <class Model, class RS, class RV, int IG>
void launch_cell_mode(cmx_ctx* c, int mode, int kind, int G, const CellArgs<RS, RV>& a)
{
    if (a.node) launch_cell_grouping<Model, RS, RV, IG, true>(c, mode, kind, G, a);
    else launch_cell_grouping<Model, RS, RV, IG, false>(c, mode, kind, G, a);
}

template <class Model, class RS, class RV, bool IMPLICIT = true>
void launch_cell_integrator(cmx_ctx* c, int integrator, int mode, int kind, int G, const CellArgs<RS, RV>& a)
{
    if (integrator == CMX_RUSH_LARSEN) launch_cell_mode<Model, RS, RV, CMX_RUSH_LARSEN>(c, mode, kind, G, a);
    else if constexpr (IMPLICIT) {
        if (integrator == CMX_BACKWARD_EULER) launch_cell_mode<Model, RS, RV, CMX_BACKWARD_EULER>(c, mode, kind, G, a);
        else if (integrator == CMX_GRL1) launch_cell_mode<Model, RS, RV, CMX_GRL1>(c, mode, kind, G, a);
        else if (integrator == CMX_GRL2) launch_cell_mode<Model, RS, RV, CMX_GRL2>(c, mode, kind, G, a);
        else launch_cell_mode<Model, RS, RV, CMX_FORWARD_EULER>(c, mode, kind, G, a);
    } else {
        launch_cell_mode<Model, RS, RV, CMX_FORWARD_EULER>(c, mode, kind, G, a);   // create admits FE / RL only
    }
}

template <class Model, class RS, class RV, class Rows>
void launch_svi_rows(cmx_ctx* c, const void* rows, const SviArgs<RS, RV>& a, int G)
{
    CMX_LAUNCH(c, CMX_K_SVI, (svi_kernel<Model, RS, RV, Rows>), G, kBlock, *static_cast<const Rows*>(rows), a);
}

template <class Model, class RS>
void launch_iionic_model(hipStream_t s, int G, int n_cells, const int32_t* node, int ps, const double* V,
                         const void* state, const void* params, const void* pdef, int n_param_planes, double cm,
                         double* out)
{
    iionic_kernel<Model, RS><<<G, kBlock, 0, s>>>(n_cells, node, ps, V, (const RS*)state, (const RS*)params,
                                                  (const RS*)pdef, n_param_planes, cm, out);
}

}  // namespace cmx
