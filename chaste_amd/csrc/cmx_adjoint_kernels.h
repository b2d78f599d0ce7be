// The adjoint's cell kernel: the Jacobian of one ICI cell step (V clamped,
// forward Euler or Rush-Larsen) by forward-mode dual numbers, contracted
// with the adjoint variables.  A template on the model, instantiated by each
// model's translation unit (cmx_model_*.hip) for its ModelOps entry.
#pragma once
#include "cmx_context.h"
#include "models/model_dual.h"

namespace cmx {

constexpr int kAdjBlock = 256;

// ---------------------------------------------------------------------------
// Cell step of one node with V clamped, as cell_step_kernel's ICI branch
// computes it for forward Euler and Rush-Larsen cells; returns Iion at the
// new state.  R = double or cmx_dualn<K>.
// ---------------------------------------------------------------------------
template <class Model, class R, int IG>
__device__ __forceinline__ R ici_cell_map(R* y, R V, const R* p, double cm, double am, const cmx_stimulus* stim,
                                          int sid, double t, double t_next, double dt_ode, int n_sub)
{
    constexpr int NS = Model::NS;
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    const R cmr = R(cm);
    y[0] = V;
#pragma unroll 1
    for (int ks = 0; ks < n_sub; ++ks) {
        const double ts = t + ks * dt_ode;
        const double h = (IG == CMX_FORWARD_EULER && ks + 1 == n_sub) ? t_next - ts : dt_ode;
        const R st = R(sid ? stimulus_at(stim[sid - 1], ts) / am : 0.0);
        R dy[NS], gi[NS], gk[NS];
        Model::template eval<R, RL>(y, V, st, p, cmr, dy, gi, gk);
        const R hh = R(h);
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            if (RL && Model::gate(s)) y[s] = gi[s] + (y[s] - gi[s]) * cmx_decay(gk[s], hh);
            else y[s] += hh * dy[s];
        }
    }
    return Model::template iion<R>(y, V, p, cmr);
}

struct CellAdjArgs {
    int n, ps, n_sub;
    double dt_ode, am, cm, c;          // c = Am Cm / dt
    StepClock clock;                   // the run's clock (t0, dt, end)
    int k;                             // the step k -> k+1
    const double* __restrict__ Vn;     // V^n
    const double* __restrict__ Yn;     // [S-1][ps] y^n
    const double* __restrict__ params; // [P][ps] or null
    const double* __restrict__ pdef;   // [P]
    int n_param_planes;
    const uint8_t* __restrict__ stim_id;
    const cmx_stimulus* __restrict__ stim;
    const double* __restrict__ lz;     // M mu
    double* __restrict__ lam_y;        // in: lambda_y^{n+1}, out: lambda_y^n   [S-1][ps]
    double* __restrict__ lam_v;        // out: lambda_V^n
    double* __restrict__ grad;         // [P][ps] accumulated dL/dp, or null
    const int32_t* __restrict__ obs_slot;  // [n]: observation column of node i, or -1
    const double* __restrict__ obs_d;  // observed values at step n, or null when n is not observed
};

// One thread per node: the Jacobian columns of (Phi, Iion) with respect to
// each state, V and each parameter, K columns per forward-mode pass,
// contracted with (lambda_y^{n+1}, -Am l_z).
template <class Model, int IG, int K>
__global__ __launch_bounds__(kAdjBlock) void cell_adjoint_kernel(CellAdjArgs a)
{
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP;
    constexpr int NPd = NP > 0 ? NP : 1;
    using Dual = cmx_dualn<K>;
    const int i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= a.n) return;
    double y0[NS], ly[NS], p0[NPd];
    y0[0] = 0.0;
    ly[0] = 0.0;
#pragma unroll
    for (int s = 1; s < NS; ++s) {
        y0[s] = a.Yn[(size_t)(s - 1) * a.ps + i];
        ly[s] = a.lam_y[(size_t)(s - 1) * a.ps + i];
    }
#pragma unroll
    for (int q = 0; q < NPd; ++q)
        p0[q] = (q < a.n_param_planes) ? a.params[(size_t)q * a.ps + i] : (NP > 0 ? a.pdef[q] : 0.0);
    const double vn = a.Vn[i];
    const int sid = a.stim_id ? (int)a.stim_id[i] : 0;
    // the step times exactly as the forward cell kernel evaluates them
    const double t = step_time(a.clock, a.k), t_next = step_next_time(a.clock, a.k);
    const double lz = a.lz[i];
    const double w_iion = -a.am * lz;
    double adj[NS + NPd];    // [1, NS): states, 0: V, NS + q: parameter q
#pragma unroll 1
    for (int k0 = 0; k0 < NS + NP; k0 += K) {
        // inputs k0 .. k0 + K - 1 (0: V, [1, NS): states, NS + q: parameters)
        Dual y[NS], p[NPd];
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            y[s] = Dual(y0[s]);
#pragma unroll
            for (int j = 0; j < K; ++j) y[s].d[j] = s == k0 + j ? 1.0 : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NPd; ++q) {
            p[q] = Dual(p0[q]);
#pragma unroll
            for (int j = 0; j < K; ++j) p[q].d[j] = NS + q == k0 + j ? 1.0 : 0.0;
        }
        Dual V(vn);
#pragma unroll
        for (int j = 0; j < K; ++j) V.d[j] = k0 + j == 0 ? 1.0 : 0.0;
        const Dual I = ici_cell_map<Model, Dual, IG>(y, V, p, a.cm, a.am, a.stim, sid, t, t_next, a.dt_ode,
                                                     a.n_sub);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double s_ = w_iion * I.d[j];
#pragma unroll
            for (int s = 1; s < NS; ++s) s_ += ly[s] * y[s].d[j];
#pragma unroll
            for (int kk = 0; kk < NS + NPd; ++kk)
                if (kk == k0 + j) adj[kk] = s_;
        }
    }
#pragma unroll
    for (int s = 1; s < NS; ++s) a.lam_y[(size_t)(s - 1) * a.ps + i] = adj[s];
    double lv = adj[0] + a.c * lz;
    if (a.obs_d) {
        const int o = a.obs_slot[i];
        if (o >= 0) lv += 2.0 * (vn - a.obs_d[o]);
    }
    a.lam_v[i] = lv;
    if (a.grad)
#pragma unroll
        for (int q = 0; q < NP; ++q) a.grad[(size_t)q * a.ps + i] += adj[NS + q];
}

template <class Model, int IG>
void launch_cell_adjoint(cmx_ctx* c, const CellAdjArgs& a)
{
    // four Jacobian columns per pass (LR91's 11 in 3 passes, TT06's 25 in 7)
    cell_adjoint_kernel<Model, IG, 4><<<(c->n + kAdjBlock - 1) / kAdjBlock, kAdjBlock, 0, c->stream>>>(a);
}

}  // namespace cmx
