// MI355X (gfx950) monodomain time-step: device kernels and the C ABI.
//
// Per PDE step (AbstractDynamicLinearPdeSolver::Solve body,
// pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:372-469):
//   1. cell_step   SolveCellSystems(V, t, t+dt) with V clamped, fused with the
//                  ICI forcing F_i = -Am Iion_i - Istim_i(t+dt)
//                  (heart/src/tissue/AbstractCardiacTissue.cpp:464-734,
//                   heart/src/solver/electrics/monodomain/MonodomainSolver.cpp:92-101)
//   2. residual    r = b - A V with b = M (Am Cm V/dt + F), A = (Am Cm/dt) M + K,
//                  evaluated as r = M F - K V (the Am Cm V/dt terms cancel
//                  exactly), with ||D^-1 b|| for the stopping test
//                  (MonodomainSolver.cpp:107 MatMult + LinearSystem::Solve)
//   3. CG          Jacobi-preconditioned CG on A d = r from d = 0 (equivalent
//                  to the reference's warm start from V^n, KSPSetInitialGuessNonzero,
//                  linalg/src/LinearSystem.cpp:833-837), stop when
//                  ||D^-1 r|| <= max(rtol ||D^-1 b||, atol)
//   4. finish      V^{n+1} = V^n + d in FP64.
// In CMX_MIXED the cell states, F and the CG vectors are FP32; V, the
// residual of step 2 (and every refinement sweep), all dot products and the
// voltage update stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cmx_internal.h"
#include "cmx_kernels.h"
#include "models/lr91.h"
#include "models/tt06_epi.h"

namespace cmx {

// ---------------------------------------------------------------------------
// Cell models
// ---------------------------------------------------------------------------
struct ModelLr91 {
    static constexpr int NS = lr91::kNumStates;
    static constexpr int NP = lr91::kNumParams;
    template <class R, bool RL>
    static __device__ __forceinline__ void eval(const R* y, R V, R st, const R* p, R cm, R* dy, R* gi, R* gk)
    {
        lr91::evaluate<R, RL>(y, V, st, p, cm, dy, gi, gk);
    }
    template <class R>
    static __device__ __forceinline__ R iion(const R* y, R V, const R* p, R cm) { return lr91::iionic<R>(y, V, p, cm); }
    static constexpr bool gate(int i) { return lr91::kIsGate[i] != 0; }
};

struct ModelTt06 {
    static constexpr int NS = tt06_epi::kNumStates;
    static constexpr int NP = tt06_epi::kNumParams;
    template <class R, bool RL>
    static __device__ __forceinline__ void eval(const R* y, R V, R st, const R* p, R cm, R* dy, R* gi, R* gk)
    {
        tt06_epi::evaluate<R, RL>(y, V, st, p, cm, dy, gi, gk);
    }
    template <class R>
    static __device__ __forceinline__ R iion(const R* y, R V, const R* p, R cm) { return tt06_epi::iionic<R>(y, V, p, cm); }
    static constexpr bool gate(int i) { return tt06_epi::kIsGate[i] != 0; }
};

template <class RS, class RV>
struct CellArgs {
    int n;
    int n_sub;                 // ODE substeps per PDE step
    double t, t_next, dt_ode;  // PDE interval and ODE step (ms)
    double am, cm;
    const double* __restrict__ V;
    RS* __restrict__ state;    // [NS-1][n]: plane s-1 holds state s (V excluded)
    const RS* __restrict__ params;  // [NP][n] or null
    const RS* __restrict__ pdef;    // [NP] defaults
    int n_param_planes;
    const uint8_t* __restrict__ stim_id;  // 0 = none, k = row k-1
    const cmx_stimulus* __restrict__ stim;
    RV* __restrict__ F;        // out: -Am Iion - Istim(t_next)
};

// One thread per node.  The state lives in registers for all substeps.
template <class Model, class RS, class RV, bool RL>
__global__ __launch_bounds__(kBlock) void cell_step_kernel(CellArgs<RS, RV> a)
{
    const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= a.n) return;
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    const double Vd = a.V[i];
    const RS V = (RS)Vd;
    RS y[NS];
    y[0] = V;
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = a.state[(size_t)(s - 1) * a.n + i];
    RS p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        p[q] = (q < a.n_param_planes) ? a.params[(size_t)q * a.n + i] : (Model::NP > 0 ? a.pdef[q] : RS(0));
    const int sid = a.stim_id ? (int)a.stim_id[i] : 0;
    const RS cm = (RS)a.cm;
    // TimeStepper(t, t_next, dt_ode): substep k starts at t + k*dt_ode
    // (global/src/TimeStepper.cpp:106-121); forward Euler takes the remaining
    // interval on the last substep, Rush-Larsen always dt_ode
    // (heart/src/odes/AbstractRushLarsenCardiacCell.cpp:106-128).
    for (int k = 0; k < a.n_sub; ++k) {
        const double ts = a.t + k * a.dt_ode;
        const double h = (!RL && k + 1 == a.n_sub) ? a.t_next - ts : a.dt_ode;
        const RS st = sid ? (RS)(stimulus_at(a.stim[sid - 1], ts) / a.am) : RS(0);
        RS dy[NS], gi[NS], gk[NS];
        Model::template eval<RS, RL>(y, V, st, p, cm, dy, gi, gk);
        const RS hh = (RS)h;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            if (RL && Model::gate(s)) y[s] = gi[s] + (y[s] - gi[s]) * exp(gk[s] * hh);
            else y[s] += hh * dy[s];
        }
    }
    const RS iion = Model::template iion<RS>(y, V, p, cm);
    const double istim = sid ? stimulus_at(a.stim[sid - 1], a.t_next) : 0.0;
    a.F[i] = (RV)(-a.am * (double)iion - istim);
#pragma unroll
    for (int s = 1; s < NS; ++s) a.state[(size_t)(s - 1) * a.n + i] = y[s];
}

// Iion from stored state (cmx_get_iionic after set_state; also used for
// the cache when no step has run).
template <class Model, class RS>
__global__ __launch_bounds__(kBlock) void iionic_kernel(int n, const double* __restrict__ V, const RS* __restrict__ state,
                                                        const RS* __restrict__ params, const RS* __restrict__ pdef,
                                                        int n_param_planes, double cm, double* __restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    constexpr int NS = Model::NS;
    constexpr int NP = Model::NP > 0 ? Model::NP : 1;
    RS y[NS];
    y[0] = (RS)V[i];
#pragma unroll
    for (int s = 1; s < NS; ++s) y[s] = state[(size_t)(s - 1) * n + i];
    RS p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        p[q] = (q < n_param_planes) ? params[(size_t)q * n + i] : (Model::NP > 0 ? pdef[q] : RS(0));
    out[i] = (double)Model::template iion<RS>(y, y[0], p, (RS)cm);
}

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

template <class RV>
struct ResidualArgs {
    int n;
    double c;                          // Am Cm / dt
    double rtol2, atol2;
    int first;                         // first sweep of the step: set tolerance
    const double* __restrict__ V;
    const double* __restrict__ Dacc;   // accumulated correction or null
    const RV* __restrict__ F;
    RV* __restrict__ r;
    RV* __restrict__ p;
    RV* __restrict__ d;
    double* __restrict__ partials;
    CgScalars* sc;
};

// r = M F - K (V + D) - c M D ; b = c M V + M F.  The stiffness term uses
// differences (V_j - V_i): K's rows sum to zero, and the differences keep the
// diffusion flux free of cancellation against the ~85 mV resting level.
template <class Rows, class RV>
__global__ __launch_bounds__(kBlock) void residual_kernel(Rows R, MatVals mv, ResidualArgs<RV> a)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = blk * kBlock + threadIdx.x;
    double part[3] = {0.0, 0.0, 0.0};
    if (i < a.n) {
        const int b = R.base(i);
        const double xi = a.V[i] + (a.Dacc ? a.Dacc[i] : 0.0);
        double mf = 0, mvv = 0, kx = 0, md = 0;
        for (int k = 0; k < R.width; ++k) {
            const int j = R.col(i, b, k);
            const int s = R.slot(i, b, k);
            const double m = mv.m64[s], kk = mv.k64[s];
            const double vj = a.V[j];
            const double dj = a.Dacc ? a.Dacc[j] : 0.0;
            mf += m * (double)a.F[j];
            mvv += m * vj;
            md += m * dj;
            kx += kk * ((vj + dj) - xi);
        }
        const double bi = a.c * mvv + mf;
        const double ri = mf - kx - a.c * md;
        const double di = mv.dinv64[R.dslot(i, b)];
        const RV rr = (RV)ri;
        const RV zz = (RV)(di * (double)rr);
        a.r[i] = rr;
        a.p[i] = zz;
        a.d[i] = RV(0);
        part[0] = (di * bi) * (di * bi);
        part[1] = (double)rr * (double)zz;
        part[2] = (double)zz * (double)zz;
    }
    double tot[3];
    if (reduce_last_block<3>(part, a.partials, &a.sc->ticket, blk, gridDim.x, tot) && threadIdx.x == 0) {
        CgScalars* sc = a.sc;
        if (a.first) {
            sc->bb = tot[0];
            sc->tol2 = fmax(a.rtol2 * tot[0], a.atol2);
        }
        sc->rz = tot[1];
        sc->zz = tot[2];
        sc->it = 0;
        sc->done = (tot[2] <= sc->tol2) ? 1 : 0;
        if (!(tot[0] == tot[0]) || !(tot[2] == tot[2])) { sc->nonfinite = 1; sc->done = 2; }
    }
}

template <class RV> __device__ __forceinline__ const RV* a_vals(const MatVals& m);
template <> __device__ __forceinline__ const double* a_vals<double>(const MatVals& m) { return m.a64; }
template <> __device__ __forceinline__ const float* a_vals<float>(const MatVals& m) { return m.a32; }
template <class RV> __device__ __forceinline__ const RV* dinv_vals(const MatVals& m);
template <> __device__ __forceinline__ const double* dinv_vals<double>(const MatVals& m) { return m.dinv64; }
template <> __device__ __forceinline__ const float* dinv_vals<float>(const MatVals& m) { return m.dinv32; }

// q = A p and (p, q); the last block forms alpha = (r, z) / (p, q).
template <class Rows, class RV>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(Rows R, MatVals mv, int n, const RV* __restrict__ p,
                                                         RV* __restrict__ q, double* __restrict__ partials, CgScalars* sc)
{
    if (sc->done) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = blk * kBlock + threadIdx.x;
    double part[1] = {0.0};
    if (i < n) {
        const RV* __restrict__ A = a_vals<RV>(mv);
        const int b = R.base(i);
        RV s = 0;
        for (int k = 0; k < R.width; ++k) s += A[R.slot(i, b, k)] * p[R.col(i, b, k)];
        q[i] = s;
        part[0] = (double)p[i] * (double)s;
    }
    double tot[1];
    if (reduce_last_block<1>(part, partials, &sc->ticket, blk, gridDim.x, tot) && threadIdx.x == 0)
        sc->alpha = sc->rz / tot[0];
}

// d += alpha p, r -= alpha q, z = D^-1 r; (r, z) and (z, z); the last block
// tests convergence and forms beta.
template <class Rows, class RV>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(Rows R, MatVals mv, int n, RV* __restrict__ d, RV* __restrict__ r,
                                                           const RV* __restrict__ p, const RV* __restrict__ q,
                                                           double* __restrict__ partials, CgScalars* sc)
{
    if (sc->done) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = blk * kBlock + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (i < n) {
        const RV alpha = (RV)sc->alpha;
        const RV di = dinv_vals<RV>(mv)[R.dslot(i, R.base(i))];
        d[i] += alpha * p[i];
        const RV ri = r[i] - alpha * q[i];
        r[i] = ri;
        const RV zi = di * ri;
        part[0] = (double)ri * (double)zi;
        part[1] = (double)zi * (double)zi;
    }
    double tot[2];
    if (reduce_last_block<2>(part, partials, &sc->ticket, blk, gridDim.x, tot) && threadIdx.x == 0) {
        sc->it += 1;
        sc->beta = tot[0] / sc->rz;
        sc->rz = tot[0];
        sc->zz = tot[1];
        if (!(tot[1] == tot[1])) { sc->nonfinite = 1; sc->done = 2; }
        else if (tot[1] <= sc->tol2) sc->done = 1;
        else if (sc->it >= sc->max_its) sc->done = 2;
    }
}

// p = D^-1 r + beta p
template <class Rows, class RV>
__global__ __launch_bounds__(kBlock) void cg_pupdate_kernel(Rows R, MatVals mv, int n, const RV* __restrict__ r,
                                                            RV* __restrict__ p, const CgScalars* sc)
{
    if (sc->done) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RV di = dinv_vals<RV>(mv)[R.dslot(i, R.base(i))];
    p[i] = di * r[i] + (RV)sc->beta * p[i];
}

// V += d (no refinement) or Dacc += d.
template <class RV>
__global__ __launch_bounds__(kBlock) void accumulate_kernel(int n, double* __restrict__ X, const RV* __restrict__ d,
                                                            CgScalars* sc)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = X[i] + (double)d[i];
    X[i] = x;
    if (!isfinite(x)) sc->nonfinite = 1;
}

__global__ __launch_bounds__(kBlock) void add_kernel(int n, double* __restrict__ V, double* __restrict__ D, CgScalars* sc)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = V[i] + D[i];
    V[i] = x;
    D[i] = 0.0;
    if (!isfinite(x)) sc->nonfinite = 1;
}

template <class T>
__global__ void fill_kernel(T* p, T v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace cmx

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
using namespace cmx;

struct cmx_ctx {
    std::string err;
    int n = 0, dim = 0, model = 0, integrator = 0, precision = 0, format = 0;
    int S = 0, P = 0, n_param_planes = 0;
    double dt = 0, dt_ode = 0, am = 0, cm = 0, rtol = 0, atol = 0;
    int max_its = 0, refine = 0, n_sub = 1, use_graph = 0;
    int width = 0, n_patterns = 0;
    long long nnz = 0;
    hipStream_t stream = nullptr;
    // device state
    double* V = nullptr;
    void* state = nullptr;       // [S-1][n] RS
    void* params = nullptr;      // [P][n] RS
    void* pdef = nullptr;        // [P] RS
    uint8_t* stim_id = nullptr;
    cmx_stimulus* stim = nullptr;
    void* F = nullptr;
    void *r = nullptr, *p = nullptr, *q = nullptr, *d = nullptr;
    double* Dacc = nullptr;
    double* iion = nullptr;      // scratch for cmx_get_iionic
    // matrix
    uint16_t* pid = nullptr;
    int32_t* off = nullptr;
    int32_t* ell_col = nullptr;
    double *m64 = nullptr, *k64 = nullptr, *a64 = nullptr, *dinv64 = nullptr;
    float *a32 = nullptr, *dinv32 = nullptr;
    double* partials = nullptr;
    CgScalars* sc = nullptr;
    CgScalars* h_sc = nullptr;   // pinned
    // host copies for cmx_get_iionic
    std::vector<cmx_stimulus> h_stim;
    std::vector<uint8_t> h_stim_id;
    double last_t_next = 0.0;
    bool stepped = false;
    // stats
    long long steps = 0, cg_total = 0, refine_total = 0;
    int last_its = 0;
    double last_residual = 0;
    double device_bytes = 0;
    std::vector<void*> allocs;
    // kernel timing (cmx_set_profiling)
    int prof_mask = 0;
    struct Pending { int kind; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double k_ms[CMX_K_COUNT] = {};
    long long k_launches[CMX_K_COUNT] = {};
};

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
    if (model == CMX_MODEL_LR91) { *S = lr91::kNumStates; *P = lr91::kNumParams; return CMX_OK; }
    if (model == CMX_MODEL_TT06_EPI) { *S = tt06_epi::kNumStates; *P = tt06_epi::kNumParams; return CMX_OK; }
    return CMX_E_ARG;
}

const double* model_init(int model) { return model == CMX_MODEL_LR91 ? lr91::kInitialState : tt06_epi::kInitialState; }
const double* model_pdef(int model) { return model == CMX_MODEL_LR91 ? lr91::kParamDefaults : tt06_epi::kParamDefaults; }

int grid_of(int n) { return (n + kBlock - 1) / kBlock; }

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

// Brackets one launch of the given kind with events when profiling it.
struct KernelTimer {
    cmx_ctx* c;
    int kind;
    hipEvent_t a = nullptr;
    KernelTimer(cmx_ctx* ctx, int k) : c(ctx), kind(k)
    {
        if (c->prof_mask & (1 << kind)) {
            a = take_event(c);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~KernelTimer()
    {
        if (!a) return;
        hipEvent_t b = take_event(c);
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({kind, a, b});
    }
};

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

// ----- typed launches -------------------------------------------------------
template <class RS, class RV>
int launch_cells(cmx_ctx* c, double t, double t_next)
{
    CellArgs<RS, RV> a;
    a.n = c->n;
    a.n_sub = c->n_sub;
    a.t = t;
    a.t_next = t_next;
    a.dt_ode = c->dt_ode;
    a.am = c->am;
    a.cm = c->cm;
    a.V = c->V;
    a.state = (RS*)c->state;
    a.params = (const RS*)c->params;
    a.pdef = (const RS*)c->pdef;
    a.n_param_planes = c->n_param_planes;
    a.stim_id = c->stim_id;
    a.stim = c->stim;
    a.F = (RV*)c->F;
    const int G = grid_of(c->n);
    const bool rl = c->integrator == CMX_RUSH_LARSEN;
    KernelTimer timer(c, CMX_K_CELL);
    if (c->model == CMX_MODEL_LR91) {
        if (rl) cell_step_kernel<ModelLr91, RS, RV, true><<<G, kBlock, 0, c->stream>>>(a);
        else cell_step_kernel<ModelLr91, RS, RV, false><<<G, kBlock, 0, c->stream>>>(a);
    } else {
        if (rl) cell_step_kernel<ModelTt06, RS, RV, true><<<G, kBlock, 0, c->stream>>>(a);
        else cell_step_kernel<ModelTt06, RS, RV, false><<<G, kBlock, 0, c->stream>>>(a);
    }
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
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

template <class Rows, class RV>
int launch_residual(cmx_ctx* c, const Rows& R, bool first, bool use_dacc)
{
    ResidualArgs<RV> a;
    a.n = c->n;
    a.c = c->am * c->cm / c->dt;
    a.rtol2 = c->rtol * c->rtol;
    a.atol2 = c->atol * c->atol;
    a.first = first ? 1 : 0;
    a.V = c->V;
    a.Dacc = use_dacc ? c->Dacc : nullptr;
    a.F = (const RV*)c->F;
    a.r = (RV*)c->r;
    a.p = (RV*)c->p;
    a.d = (RV*)c->d;
    a.partials = c->partials;
    a.sc = c->sc;
    KernelTimer timer(c, CMX_K_RESIDUAL);
    residual_kernel<Rows, RV><<<grid_of(c->n), kBlock, 0, c->stream>>>(R, mat_vals(c), a);
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

template <class Rows, class RV>
int launch_cg_iters(cmx_ctx* c, const Rows& R, int iters)
{
    const int G = grid_of(c->n);
    const MatVals mv = mat_vals(c);
    for (int k = 0; k < iters; ++k) {
        {
            KernelTimer timer(c, CMX_K_SPMV);
            cg_spmv_kernel<Rows, RV><<<G, kBlock, 0, c->stream>>>(R, mv, c->n, (const RV*)c->p, (RV*)c->q, c->partials, c->sc);
        }
        {
            KernelTimer timer(c, CMX_K_UPDATE);
            cg_update_kernel<Rows, RV><<<G, kBlock, 0, c->stream>>>(R, mv, c->n, (RV*)c->d, (RV*)c->r, (const RV*)c->p,
                                                                   (const RV*)c->q, c->partials, c->sc);
        }
        {
            KernelTimer timer(c, CMX_K_PUPDATE);
            cg_pupdate_kernel<Rows, RV><<<G, kBlock, 0, c->stream>>>(R, mv, c->n, (const RV*)c->r, (RV*)c->p, c->sc);
        }
    }
    CMX_HIP(c, hipGetLastError());
    return CMX_OK;
}

int read_scalars(cmx_ctx* c)
{
    CMX_HIP(c, hipMemcpyAsync(c->h_sc, c->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->pending.empty()) collect_times(c);
    return CMX_OK;
}

// One linear solve on the residual already in r/p (d = 0): enqueue chunks of
// iterations, each a no-op once converged, until the control block says done.
template <class Rows, class RV>
int solve(cmx_ctx* c, const Rows& R)
{
    // The first chunk is sized from the previous step's count, so a step
    // normally costs one host synchronisation.
    int chunk = std::min(std::max(c->last_its + 1, 2), c->max_its);
    int launched = 0;
    int rc;
    do {
        rc = launch_cg_iters<Rows, RV>(c, R, chunk);
        if (rc) return rc;
        launched += chunk;
        rc = read_scalars(c);
        if (rc) return rc;
        chunk = std::min(4, std::max(c->max_its - launched, 1));
    } while (c->h_sc->done == 0);
    if (c->h_sc->nonfinite) return fail(c, CMX_E_NONFINITE, "non-finite residual in the diffusion solve");
    if (c->h_sc->done == 2)
        return fail(c, CMX_E_DIVERGED, "CG reached the iteration limit (" + std::to_string(c->max_its) +
                                           ") without converging: KSP_DIVERGED_ITS");
    return CMX_OK;
}

template <class Rows, class RS, class RV>
int step_once(cmx_ctx* c, const Rows& R, double t, double t_next)
{
    int rc = launch_cells<RS, RV>(c, t, t_next);
    if (rc) return rc;
    const int G = grid_of(c->n);
    const int sweeps = (c->precision == CMX_MIXED) ? std::max(c->refine, 0) : 0;
    rc = launch_residual<Rows, RV>(c, R, true, false);
    if (rc) return rc;
    rc = solve<Rows, RV>(c, R);
    if (rc) return rc;
    int its = c->h_sc->it;
    if (sweeps == 0) {
        KernelTimer timer(c, CMX_K_FINISH);
        accumulate_kernel<RV><<<G, kBlock, 0, c->stream>>>(c->n, c->V, (const RV*)c->d, c->sc);
    } else {
        // FP64 refinement: true residual of V^n + Dacc against the step's b.
        accumulate_kernel<RV><<<G, kBlock, 0, c->stream>>>(c->n, c->Dacc, (const RV*)c->d, c->sc);
        for (int s = 0; s < sweeps; ++s) {
            rc = launch_residual<Rows, RV>(c, R, false, true);
            if (rc) return rc;
            ++c->refine_total;
            rc = read_scalars(c);
            if (rc) return rc;
            if (c->h_sc->done) break;   // true residual already below the threshold
            rc = solve<Rows, RV>(c, R);
            if (rc) return rc;
            its += c->h_sc->it;
            accumulate_kernel<RV><<<G, kBlock, 0, c->stream>>>(c->n, c->Dacc, (const RV*)c->d, c->sc);
        }
        add_kernel<<<G, kBlock, 0, c->stream>>>(c->n, c->V, c->Dacc, c->sc);
    }
    CMX_HIP(c, hipGetLastError());
    c->last_its = its;
    c->cg_total += its;
    c->last_residual = std::sqrt(std::max(c->h_sc->zz, 0.0));
    c->steps += 1;
    c->stepped = true;
    c->last_t_next = t_next;
    return CMX_OK;
}

template <class RS, class RV>
int step_typed(cmx_ctx* c, double t, double t_next)
{
    if (c->format == CMX_FMT_STENCIL) {
        StencilRows R{c->pid, c->off, c->width};
        return step_once<StencilRows, RS, RV>(c, R, t, t_next);
    }
    EllRows R{c->ell_col, c->width, c->n};
    return step_once<EllRows, RS, RV>(c, R, t, t_next);
}

template <class RS>
int to_device_planes(cmx_ctx* c, void* dst, const double* src, size_t count)
{
    std::vector<RS> tmp(src, src + count);
    CMX_HIP(c, hipMemcpy(dst, tmp.data(), count * sizeof(RS), hipMemcpyHostToDevice));
    return CMX_OK;
}

template <class RS>
int from_device_planes(cmx_ctx* c, double* dst, const void* src, size_t count)
{
    std::vector<RS> tmp(count);
    CMX_HIP(c, hipMemcpy(tmp.data(), src, count * sizeof(RS), hipMemcpyDeviceToHost));
    std::copy(tmp.begin(), tmp.end(), dst);
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

int create_impl(const cmx_mesh* mesh, const cmx_tissue* tissue, const cmx_opts* opts, cmx_ctx* c)
{
    if (!mesh || !tissue || !opts) return fail(c, CMX_E_ARG, "null mesh, tissue or opts");
    if (model_sizes(tissue->model, &c->S, &c->P)) return fail(c, CMX_E_ARG, "unknown cell model");
    if (tissue->integrator != CMX_FORWARD_EULER && tissue->integrator != CMX_RUSH_LARSEN)
        return fail(c, CMX_E_ARG, "unknown integrator");
    if (opts->precision != CMX_FP64 && opts->precision != CMX_MIXED) return fail(c, CMX_E_ARG, "unknown precision");
    if (!(opts->dt_pde > 0) || !(opts->dt_ode > 0) || opts->dt_ode > opts->dt_pde * (1 + 1e-12))
        return fail(c, CMX_E_ARG, "need 0 < dt_ode <= dt_pde");
    const double ratio = opts->dt_pde / opts->dt_ode;
    const int n_sub = (int)std::floor(ratio + 0.5);
    if (std::fabs(ratio - n_sub) > 1e-9 * ratio)
        return fail(c, CMX_E_ARG, "dt_ode must divide dt_pde");
    if (!(opts->am > 0) || !(opts->cm > 0)) return fail(c, CMX_E_ARG, "Am and Cm must be positive");
    if (opts->ksp_max_its <= 0) return fail(c, CMX_E_ARG, "ksp_max_its must be positive");
    if (tissue->n_params < 0 || tissue->n_params > c->P || (tissue->n_params > 0 && !tissue->params))
        return fail(c, CMX_E_ARG, "bad per-node parameter planes");
    if (tissue->n_stimuli < 0 || tissue->n_stimuli > 255 || (tissue->n_stimuli > 0 && !tissue->stimuli))
        return fail(c, CMX_E_ARG, "bad stimulus table (at most 255 rows)");
    c->n = mesh->n_nodes;
    c->dim = mesh->dim;
    c->model = tissue->model;
    c->integrator = tissue->integrator;
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
    c->use_graph = opts->use_graph;
    c->n_param_planes = tissue->n_params;

    HostCsr csr;
    std::string err;
    if (!assemble_p1(*mesh, opts->mass_lumping != 0, csr, err)) return fail(c, CMX_E_ARG, err);
    c->nnz = csr.row_ptr[c->n];
    c->width = csr.max_row;
    const int n = c->n;
    const double cA = c->am * c->cm / c->dt;

    CMX_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    int fmt = opts->format;
    StencilTable st;
    if (fmt == CMX_FMT_AUTO || fmt == CMX_FMT_STENCIL) {
        if (build_stencil(csr, 4096, st)) fmt = CMX_FMT_STENCIL;
        else if (fmt == CMX_FMT_STENCIL) return fail(c, CMX_E_ARG, "mesh has too many distinct row patterns for the stencil format");
        else fmt = CMX_FMT_ELL;
    }
    c->format = fmt;
    int rc;
    if (fmt == CMX_FMT_STENCIL) {
        const int W = st.width, NPt = st.n_patterns;
        c->n_patterns = NPt;
        std::vector<double> a64(st.m.size()), dinv64(NPt);
        std::vector<float> a32(st.m.size()), dinv32(NPt);
        for (size_t s = 0; s < st.m.size(); ++s) {
            a64[s] = (1.0 / c->dt) * (c->am * c->cm * st.m[s]) + st.k[s];
            a32[s] = (float)a64[s];
        }
        for (int q = 0; q < NPt; ++q) {
            double diag = 0;
            for (int k = 0; k < W; ++k)
                if (st.off[(size_t)q * W + k] == 0) diag += a64[(size_t)q * W + k];
            dinv64[q] = 1.0 / diag;
            dinv32[q] = (float)dinv64[q];
        }
        if ((rc = upload(c, &c->pid, st.pid.data(), n))) return rc;
        if ((rc = upload(c, &c->off, st.off.data(), st.off.size()))) return rc;
        if ((rc = upload(c, &c->m64, st.m.data(), st.m.size()))) return rc;
        if ((rc = upload(c, &c->k64, st.k.data(), st.k.size()))) return rc;
        if ((rc = upload(c, &c->a64, a64.data(), a64.size()))) return rc;
        if ((rc = upload(c, &c->a32, a32.data(), a32.size()))) return rc;
        if ((rc = upload(c, &c->dinv64, dinv64.data(), NPt))) return rc;
        if ((rc = upload(c, &c->dinv32, dinv32.data(), NPt))) return rc;
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
                if (csr.col[p] == i) dinv64[i] = 1.0 / ((1.0 / c->dt) * (c->am * c->cm * csr.mass[p]) + csr.stiff[p]);
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
    (void)cA;

    // Cells: default initial state (AbstractCardiacCell::Init ->
    // ResetToInitialConditions), V from the model's initial voltage.
    const bool mixed = c->precision == CMX_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    const size_t rv = rs;
    const double* y0 = model_init(c->model);
    std::vector<double> planes((size_t)(c->S - 1) * n);
    for (int s = 1; s < c->S; ++s) std::fill(planes.begin() + (size_t)(s - 1) * n, planes.begin() + (size_t)s * n, y0[s]);
    std::vector<double> V0(n, y0[0]);
    if ((rc = upload(c, &c->V, V0.data(), n))) return rc;
    if ((rc = dalloc(c, (char**)&c->state, planes.size() * rs))) return rc;
    if ((rc = mixed ? to_device_planes<float>(c, c->state, planes.data(), planes.size())
                    : to_device_planes<double>(c, c->state, planes.data(), planes.size())))
        return rc;
    const int P = std::max(c->P, 1);
    std::vector<double> pdef(model_pdef(c->model), model_pdef(c->model) + P);
    if ((rc = dalloc(c, (char**)&c->pdef, P * rs))) return rc;
    if ((rc = mixed ? to_device_planes<float>(c, c->pdef, pdef.data(), P) : to_device_planes<double>(c, c->pdef, pdef.data(), P)))
        return rc;
    if (c->n_param_planes > 0) {
        const size_t np = (size_t)c->n_param_planes * n;
        if ((rc = dalloc(c, (char**)&c->params, np * rs))) return rc;
        if ((rc = mixed ? to_device_planes<float>(c, c->params, tissue->params, np)
                        : to_device_planes<double>(c, c->params, tissue->params, np)))
            return rc;
    }
    // Stimulus table (durations padded as the reference's constructors do).
    c->h_stim.resize(tissue->n_stimuli);
    for (int k = 0; k < tissue->n_stimuli; ++k) {
        cmx_stimulus s = tissue->stimuli[k];
        if (s.period > 0) {
            if (s.duration > s.period) return fail(c, CMX_E_ARG, "stimulus duration exceeds its period");
            s.duration += s.period * DBL_EPSILON;
        } else {
            s.duration += (s.duration + s.start) * DBL_EPSILON;
        }
        if (!(s.stop > 0)) s.stop = DBL_MAX;
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
    if ((rc = dalloc(c, (char**)&c->F, n * rv))) return rc;
    if ((rc = dalloc(c, (char**)&c->r, n * rv))) return rc;
    if ((rc = dalloc(c, (char**)&c->p, n * rv))) return rc;
    if ((rc = dalloc(c, (char**)&c->q, n * rv))) return rc;
    if ((rc = dalloc(c, (char**)&c->d, n * rv))) return rc;
    if ((rc = dalloc(c, &c->Dacc, n))) return rc;
    CMX_HIP(c, hipMemset(c->Dacc, 0, n * sizeof(double)));
    CMX_HIP(c, hipMemset(c->F, 0, n * rv));
    if ((rc = dalloc(c, &c->partials, 3 * (size_t)grid_of(n)))) return rc;
    if ((rc = dalloc(c, &c->sc, 1))) return rc;
    CgScalars z{};
    z.max_its = c->max_its;
    CMX_HIP(c, hipMemcpy(c->sc, &z, sizeof(z), hipMemcpyHostToDevice));
    CMX_HIP(c, hipHostMalloc((void**)&c->h_sc, sizeof(CgScalars), hipHostMallocDefault));
    *c->h_sc = z;
    CMX_HIP(c, hipDeviceSynchronize());
    return CMX_OK;
}

void release(cmx_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_times(c);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    c->h_sc = nullptr;
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
    // TimeStepper(t0, t0 + n dt, dt): t_k = t0 + k dt, the end snapped
    // (global/src/TimeStepper.cpp:106-121).
    const double t_end = t0 + n_steps * c->dt;
    const double eps = t_end > 1.0 ? DBL_EPSILON * t_end : DBL_EPSILON;
    for (int k = 0; k < n_steps; ++k) {
        const double t = t0 + k * c->dt;
        double tn = t0 + (k + 1) * c->dt;
        if (t_end - tn <= eps) tn = t_end;
        const int rc = (c->precision == CMX_MIXED) ? step_typed<float, float>(c, t, tn) : step_typed<double, double>(c, t, tn);
        if (rc) return rc;
    }
    int rc = read_scalars(c);
    if (rc) return rc;
    if (c->h_sc->nonfinite) return fail(c, CMX_E_NONFINITE, "non-finite transmembrane potential (cell ODE or solve blew up)");
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
    return CMX_OK;
}

int cmx_get_state(cmx_ctx* c, double* states)
{
    if (!c || !states) return fail(c, CMX_E_ARG, "null argument");
    int rc = cmx_get_voltage(c, states);
    if (rc) return rc;
    const size_t cnt = (size_t)(c->S - 1) * c->n;
    return c->precision == CMX_MIXED ? from_device_planes<float>(c, states + c->n, c->state, cnt)
                                     : from_device_planes<double>(c, states + c->n, c->state, cnt);
}

int cmx_set_state(cmx_ctx* c, const double* states)
{
    if (!c || !states) return fail(c, CMX_E_ARG, "null argument");
    int rc = cmx_set_voltage(c, states);
    if (rc) return rc;
    const size_t cnt = (size_t)(c->S - 1) * c->n;
    return c->precision == CMX_MIXED ? to_device_planes<float>(c, c->state, states + c->n, cnt)
                                     : to_device_planes<double>(c, c->state, states + c->n, cnt);
}

int cmx_get_iionic(cmx_ctx* c, double* out)
{
    if (!c || !out) return fail(c, CMX_E_ARG, "null argument");
    if (!c->stepped) {
        // No step yet: evaluate at the current state.
        if (!c->iion) {
            int rc = dalloc(c, &c->iion, c->n);
            if (rc) return rc;
        }
        const int G = grid_of(c->n);
        if (c->precision == CMX_MIXED) {
            if (c->model == CMX_MODEL_LR91)
                iionic_kernel<ModelLr91, float><<<G, kBlock, 0, c->stream>>>(c->n, c->V, (const float*)c->state, (const float*)c->params, (const float*)c->pdef, c->n_param_planes, c->cm, c->iion);
            else
                iionic_kernel<ModelTt06, float><<<G, kBlock, 0, c->stream>>>(c->n, c->V, (const float*)c->state, (const float*)c->params, (const float*)c->pdef, c->n_param_planes, c->cm, c->iion);
        } else {
            if (c->model == CMX_MODEL_LR91)
                iionic_kernel<ModelLr91, double><<<G, kBlock, 0, c->stream>>>(c->n, c->V, (const double*)c->state, (const double*)c->params, (const double*)c->pdef, c->n_param_planes, c->cm, c->iion);
            else
                iionic_kernel<ModelTt06, double><<<G, kBlock, 0, c->stream>>>(c->n, c->V, (const double*)c->state, (const double*)c->params, (const double*)c->pdef, c->n_param_planes, c->cm, c->iion);
        }
        CMX_HIP(c, hipGetLastError());
        CMX_HIP(c, hipStreamSynchronize(c->stream));
        CMX_HIP(c, hipMemcpy(out, c->iion, c->n * sizeof(double), hipMemcpyDeviceToHost));
        return CMX_OK;
    }
    // After a step the cache is recovered from F = -Am Iion - Istim(t_next).
    std::vector<double> F(c->n);
    CMX_HIP(c, hipStreamSynchronize(c->stream));
    int rc = c->precision == CMX_MIXED ? from_device_planes<float>(c, F.data(), c->F, c->n)
                                       : from_device_planes<double>(c, F.data(), c->F, c->n);
    if (rc) return rc;
    for (int i = 0; i < c->n; ++i) {
        const int sid = c->h_stim_id[i];
        const double is = sid ? host_stimulus(c->h_stim[sid - 1], c->last_t_next) : 0.0;
        out[i] = -(F[i] + is) / c->am;
    }
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

int cmx_model_info(int32_t model, int32_t* n_states, int32_t* n_params)
{
    int S, P;
    if (model_sizes(model, &S, &P)) return fail(nullptr, CMX_E_ARG, "unknown cell model");
    if (n_states) *n_states = S;
    if (n_params) *n_params = P;
    return CMX_OK;
}

int cmx_model_state_name(int32_t model, int32_t index, const char** name)
{
    int S, P;
    if (model_sizes(model, &S, &P) || index < 0 || index >= S || !name) return fail(nullptr, CMX_E_ARG, "bad model/index");
    *name = model == CMX_MODEL_LR91 ? lr91::kStateNames[index] : tt06_epi::kStateNames[index];
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
