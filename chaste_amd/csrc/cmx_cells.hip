// Batched single cells (cmx_cells_run): many independent cells, each
// integrated from t = 0 to its own end time with its own parameters and
// stimulus, as AbstractCardiacCell::Compute integrates one cell
// (heart/src/odes/AbstractCardiacCell.cpp:64-88 with
// ode/src/solver/AbstractOneStepIvpOdeSolver.cpp:93-153 for forward Euler,
// heart/src/odes/AbstractRushLarsenCardiacCell.cpp:57-104 for Rush-Larsen):
// every rate at the step's start, V by forward Euler, the area stimulus of
// the cell's protocol at the step's start time.  Each cell's voltage trace is
// summarised on the fly as CellProperties does for action potentials
// (heart/src/postprocessing/CellProperties.cpp:54-200): onsets where V rises
// above the threshold (linear interpolation), peaks, and the threshold
// crossing that ends each action potential.
//
// One thread per cell, the state in registers across the steps of a launch.
// Cells are sorted by run length so that the lanes of a wavefront finish
// together; launches of a fixed number of steps follow one another on the
// stream (no host synchronisation), each wavefront leaving as soon as its
// cells are done.  The launches are VALU-bound: a cell step is several
// hundred floating-point operations and ~30 exponentials against ~200 bytes
// of state moved per launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "cmx_context.h"

namespace cmx {
namespace {

thread_local std::string g_cells_error;

int cfail(int code, const std::string& msg)
{
    g_cells_error = msg;
    return code;
}

#define CELLS_HIP(call)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return cfail(CMX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class RS>
struct CellsArgs {
    int n, ps;
    double dt, cm, thr;
    const int32_t* __restrict__ n_steps;   // [n] steps of each cell (sorted, longest first)
    const cmx_stimulus* __restrict__ stim; // [n] padded protocols, or null
    double* __restrict__ state;            // [S-1][ps]
    double* __restrict__ V;                // [n]
    const RS* __restrict__ params;         // [P][ps] or null
    const RS* __restrict__ pdef;           // [P]
    int n_param_planes;
    // CellProperties summary: [ps] each
    int32_t* __restrict__ phase;           // 0 below threshold, 1 above
    int32_t* __restrict__ ap_count;        // completed action potentials
    double* __restrict__ onset;            // last onset
    double* __restrict__ peak;             // running peak of the current action potential
    double* __restrict__ last_peak;        // peak of the last completed one
    double* __restrict__ end;              // threshold crossing that ended it
};

template <class Model, class RS, int IG>
__global__ __launch_bounds__(kBlock) void cells_kernel(CellsArgs<RS> a, int k0, int k1)
{
    constexpr int NS = Model::NS;
    constexpr int NPd = Model::NP > 0 ? Model::NP : 1;
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int kend = min(k1, a.n_steps[i]);
    if (k0 >= kend) return;
    // Rates in RS; the states that integrate small increments over long runs
    // (V and every non-gate: concentrations) accumulate in FP64 whatever RS
    // is -- an FP32 K_i of ~138 mM cannot take a per-step change of ~1e-6 mM.
    // Gates (bounded, Rush-Larsen or Euler) stay in RS.
    RS y[NS], p[NPd];
    double yd[NS];
#pragma unroll
    for (int s = 1; s < NS; ++s) {
        yd[s] = a.state[(size_t)(s - 1) * a.ps + i];
        y[s] = (RS)yd[s];
    }
#pragma unroll
    for (int q = 0; q < NPd; ++q)
        p[q] = (q < a.n_param_planes) ? a.params[(size_t)q * a.ps + i] : (Model::NP > 0 ? a.pdef[q] : RS(0));
    double vd = a.V[i];
    cmx_stimulus st{};
    if (a.stim) st = a.stim[i];
    int phase = a.phase[i], count = a.ap_count[i];
    double onset = a.onset[i], peak = a.peak[i], last_peak = a.last_peak[i], end = a.end[i];
    const RS cm = (RS)a.cm, h = (RS)a.dt;
    const double thr = a.thr;
    for (int k = k0; k < kend; ++k) {
        const double t = k * a.dt;
        const RS is = (RS)(a.stim ? stimulus_at(st, t) : 0.0);
        const RS V = (RS)vd;
        y[0] = V;
        RS dy[NS], gi[NS], gk[NS];
        Model::template eval<RS, RL>(y, V, is, p, cm, dy, gi, gk);
        const RS dv = Model::template dvdt<RS>(y, V, is, p, cm);
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            if (Model::gate(s)) {
                if (RL) y[s] = gi[s] + (y[s] - gi[s]) * cmx_exp(gk[s] * h);
                else y[s] += h * dy[s];
            } else {
                yd[s] += a.dt * (double)dy[s];
                y[s] = (RS)yd[s];
            }
        }
        const double v0 = vd;
        vd += a.dt * (double)dv;
        // CellProperties: onset when V rises above the threshold, the peak
        // while above, the end when V falls below it again
        const double t1 = (k + 1) * a.dt;
        if (phase == 0 && vd > thr && v0 <= thr) {
            onset = t + (t1 - t) / (vd - v0) * (thr - v0);
            phase = 1;
        }
        if (phase == 1) {
            if (vd > peak) peak = vd;
            if (vd < thr && v0 >= thr) {
                last_peak = peak;
                peak = thr;
                end = t + (t1 - t) / (vd - v0) * (thr - v0);
                ++count;
                phase = 0;
            }
        }
    }
#pragma unroll
    for (int s = 1; s < NS; ++s) a.state[(size_t)(s - 1) * a.ps + i] = Model::gate(s) ? (double)y[s] : yd[s];
    a.V[i] = vd;
    a.phase[i] = phase;
    a.ap_count[i] = count;
    a.onset[i] = onset;
    a.peak[i] = peak;
    a.last_peak[i] = last_peak;
    a.end[i] = end;
}

struct Buffers {
    std::vector<void*> ptrs;
    ~Buffers()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** out, size_t count)
    {
        void* p = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return cfail(CMX_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
        ptrs.push_back(p);
        *out = (T*)p;
        return CMX_OK;
    }
    template <class T>
    int upload(T** out, const std::vector<T>& host)
    {
        int rc = alloc(out, host.size());
        if (rc) return rc;
        if (!host.empty()) CELLS_HIP(hipMemcpy(*out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
        return CMX_OK;
    }
};

template <class Model, class RS, int IG>
void launch_cells(const CellsArgs<RS>& a, int k0, int k1, hipStream_t s)
{
    cells_kernel<Model, RS, IG><<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a, k0, k1);
}

template <class RS>
int run_cells(const cmx_cells& c, int S, int P, const double* y0, double* y_out, double* props, cmx_cells_stats* stats)
{
    const int n = c.n_cells;
    const int ps = (n + 63) / 64 * 64;
    // per-cell step counts, cells sorted longest first
    std::vector<int32_t> steps(n);
    for (int i = 0; i < n; ++i) {
        const double r = c.t_end[i] / c.dt;
        const long long k = std::llround(r);
        if (!(c.t_end[i] >= 0) || std::fabs(r - (double)k) > 1e-9 * std::max(1.0, r) || k > INT32_MAX)
            return cfail(CMX_E_ARG, "every t_end must be a non-negative whole number of dt");
        steps[i] = (int32_t)k;
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return steps[a] > steps[b]; });
    std::vector<int32_t> s_steps(n);
    std::vector<cmx_stimulus> s_stim;
    std::vector<double> s_state((size_t)(S - 1) * ps, 0.0);
    std::vector<RS> s_params((size_t)std::max(c.n_params, 0) * ps, RS(0));
    std::vector<double> s_v(n);
    const double* yi = y0 ? y0 : nullptr;
    std::vector<double> def(S);
    if (!yi) {
        const double* init = c.model == CMX_MODEL_LR91 ? lr91::kInitialState : tt06_epi::kInitialState;
        std::copy(init, init + S, def.begin());
    }
    if (c.stimuli) s_stim.resize(n);
    for (int j = 0; j < n; ++j) {
        const int i = order[j];
        s_steps[j] = steps[i];
        s_v[j] = yi ? yi[i] : def[0];
        for (int s = 1; s < S; ++s) s_state[(size_t)(s - 1) * ps + j] = yi ? yi[(size_t)s * n + i] : def[s];
        for (int q = 0; q < c.n_params; ++q) s_params[(size_t)q * ps + j] = (RS)c.params[(size_t)q * n + i];
        if (c.stimuli) {
            cmx_stimulus st = c.stimuli[i];
            if (!pad_stimulus(st)) return cfail(CMX_E_ARG, "stimulus duration exceeds its period");
            s_stim[j] = st;
        }
    }
    std::vector<RS> pdef(std::max(P, 1), RS(0));
    const double* pd = c.model == CMX_MODEL_LR91 ? lr91::kParamDefaults : tt06_epi::kParamDefaults;
    for (int q = 0; q < P; ++q) pdef[q] = (RS)pd[q];

    Buffers B;
    CellsArgs<RS> a{};
    int rc;
    int32_t* d_steps;
    cmx_stimulus* d_stim = nullptr;
    RS *d_params = nullptr, *d_pdef;
    double *d_state, *d_v;
    if ((rc = B.upload(&d_steps, s_steps))) return rc;
    if (c.stimuli && (rc = B.upload(&d_stim, s_stim))) return rc;
    if ((rc = B.upload(&d_state, s_state))) return rc;
    if (c.n_params > 0 && (rc = B.upload(&d_params, s_params))) return rc;
    if ((rc = B.upload(&d_pdef, pdef))) return rc;
    if ((rc = B.upload(&d_v, s_v))) return rc;
    std::vector<int32_t> zeros_i(ps, 0);
    std::vector<double> minus1(ps, -1.0), lowest(ps, -DBL_MAX);
    int32_t *d_phase, *d_count;
    double *d_onset, *d_peak, *d_last_peak, *d_end;
    if ((rc = B.upload(&d_phase, zeros_i))) return rc;
    if ((rc = B.upload(&d_count, zeros_i))) return rc;
    if ((rc = B.upload(&d_onset, minus1))) return rc;
    if ((rc = B.upload(&d_peak, lowest))) return rc;
    if ((rc = B.upload(&d_last_peak, minus1))) return rc;
    if ((rc = B.upload(&d_end, minus1))) return rc;
    a.phase = d_phase;
    a.ap_count = d_count;
    a.onset = d_onset;
    a.peak = d_peak;
    a.last_peak = d_last_peak;
    a.end = d_end;
    a.n = n;
    a.ps = ps;
    a.dt = c.dt;
    a.cm = 1.0;
    a.thr = c.threshold;
    a.n_steps = d_steps;
    a.stim = d_stim;
    a.state = d_state;
    a.V = d_v;
    a.params = d_params;
    a.pdef = d_pdef;
    a.n_param_planes = c.n_params;

    hipStream_t stream = nullptr;
    CELLS_HIP(hipStreamCreate(&stream));
    struct StreamFree {
        hipStream_t s;
        ~StreamFree() { (void)hipStreamDestroy(s); }
    } sf{stream};
    hipEvent_t e0, e1;
    CELLS_HIP(hipEventCreate(&e0));
    CELLS_HIP(hipEventCreate(&e1));
    struct EventFree {
        hipEvent_t a, b;
        ~EventFree()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ef{e0, e1};
    const int longest = n > 0 ? s_steps[0] : 0;
    const int chunk = std::max(1, c.steps_per_launch > 0 ? c.steps_per_launch : 2000);
    int launches = 0;
    CELLS_HIP(hipEventRecord(e0, stream));
    for (int k0 = 0; k0 < longest; k0 += chunk) {
        const int k1 = std::min(longest, k0 + chunk);
        // only the cells still running: the sorted prefix with n_steps > k0
        CellsArgs<RS> ak = a;
        // (counts sorted descending: lower_bound with greater<> finds the first <= k0)
        ak.n = (int)(std::lower_bound(s_steps.begin(), s_steps.end(), k0, std::greater<int32_t>()) - s_steps.begin());
        if (ak.n <= 0) break;
        const bool rl = c.integrator == CMX_RUSH_LARSEN;
        if (c.model == CMX_MODEL_LR91) {
            if (rl) launch_cells<ModelLr91, RS, CMX_RUSH_LARSEN>(ak, k0, k1, stream);
            else launch_cells<ModelLr91, RS, CMX_FORWARD_EULER>(ak, k0, k1, stream);
        } else {
            if (rl) launch_cells<ModelTt06, RS, CMX_RUSH_LARSEN>(ak, k0, k1, stream);
            else launch_cells<ModelTt06, RS, CMX_FORWARD_EULER>(ak, k0, k1, stream);
        }
        CELLS_HIP(hipGetLastError());
        ++launches;
    }
    CELLS_HIP(hipEventRecord(e1, stream));
    CELLS_HIP(hipStreamSynchronize(stream));
    float ms = 0.f;
    CELLS_HIP(hipEventElapsedTime(&ms, e0, e1));

    // back to the caller's order
    CELLS_HIP(hipMemcpy(s_state.data(), d_state, s_state.size() * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(s_v.data(), d_v, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> cnt(ps);
    std::vector<double> on(ps), lp(ps), en(ps);
    CELLS_HIP(hipMemcpy(cnt.data(), a.ap_count, ps * sizeof(int32_t), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(on.data(), a.onset, ps * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(lp.data(), a.last_peak, ps * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(en.data(), a.end, ps * sizeof(double), hipMemcpyDeviceToHost));
    long long cell_steps = 0;
    for (int j = 0; j < n; ++j) {
        const int i = order[j];
        cell_steps += s_steps[j];
        if (y_out) {
            y_out[i] = s_v[j];
            for (int s = 1; s < S; ++s) y_out[(size_t)s * n + i] = s_state[(size_t)(s - 1) * ps + j];
        }
        if (props) {
            props[i] = (double)cnt[j];
            props[(size_t)n + i] = on[j];
            props[2 * (size_t)n + i] = lp[j];
            props[3 * (size_t)n + i] = en[j];
        }
    }
    if (stats) {
        stats->kernel_ms = ms;
        stats->cell_steps = cell_steps;
        stats->launches = launches;
        stats->block = kBlock;
    }
    return CMX_OK;
}

}  // namespace
}  // namespace cmx

using namespace cmx;

extern "C" {

int cmx_cells_run(const cmx_cells* c, const double* y0, double* y_out, double* props, cmx_cells_stats* stats)
{
    if (!c) return cfail(CMX_E_ARG, "null description");
    int S, P;
    if (c->model == CMX_MODEL_LR91) { S = lr91::kNumStates; P = lr91::kNumParams; }
    else if (c->model == CMX_MODEL_TT06_EPI) { S = tt06_epi::kNumStates; P = tt06_epi::kNumParams; }
    else return cfail(CMX_E_ARG, "unknown cell model");
    if (c->integrator != CMX_FORWARD_EULER && c->integrator != CMX_RUSH_LARSEN)
        return cfail(CMX_E_ARG, "batched cells support forward-Euler and Rush-Larsen integration");
    if (c->precision != CMX_FP64 && c->precision != CMX_MIXED) return cfail(CMX_E_ARG, "unknown precision");
    if (c->n_cells < 0 || (c->n_cells > 0 && !c->t_end)) return cfail(CMX_E_ARG, "n_cells and t_end are required");
    if (!(c->dt > 0)) return cfail(CMX_E_ARG, "dt must be positive");
    if (c->n_params < 0 || c->n_params > P || (c->n_params > 0 && !c->params))
        return cfail(CMX_E_ARG, "n_params must be in [0, the model's parameter count] with params given");
    if (c->n_cells == 0) return CMX_OK;
    return c->precision == CMX_MIXED ? run_cells<float>(*c, S, P, y0, y_out, props, stats)
                                     : run_cells<double>(*c, S, P, y0, y_out, props, stats);
}

const char* cmx_cells_last_error(void) { return g_cells_error.c_str(); }

}  // extern "C"
