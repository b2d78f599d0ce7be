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
#include <climits>
#include <cmath>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "cmx_models.h"

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
        ModelMeta mm{};
        model_meta(c.model, &mm);
        std::copy(mm.init, mm.init + S, def.begin());
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
    ModelMeta mp{};
    model_meta(c.model, &mp);
    for (int q = 0; q < P; ++q) pdef[q] = (RS)mp.pdef[q];

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
    int32_t* d_miss = nullptr;
    if (c.lookup_tables) {
        if (!std::is_same<RS, float>::value || c.integrator != CMX_RUSH_LARSEN)
            return cfail(CMX_E_ARG, "lookup tables need mixed precision and Rush-Larsen cells");
        std::vector<float> table;
        int rows = 0;
        const char* err = gate_table(c.model, c.dt, 1.0, table, &rows);
        if (err) return cfail(CMX_E_ARG, err);
        float* d_lut;
        if ((rc = B.upload(&d_lut, table))) return rc;
        if ((rc = B.upload(&d_miss, std::vector<int32_t>(1, 0)))) return rc;
        a.lut = d_lut;
        a.lut_vmin = (float)kLutVmin;
        a.lut_inv_dv = (float)(1.0 / kLutDv);
        a.lut_rows = rows;
        a.lut_miss = d_miss;
    }

    // V trace: device samples in the sorted order, NaN until written
    double* d_trace = nullptr;
    size_t n_samples = 0;
    if (c.trace_every > 0) {
        n_samples = (size_t)(n > 0 ? s_steps[0] : 0) / (size_t)c.trace_every + 1;
        if (n_samples * (size_t)n > ((size_t)1 << 31))
            return cfail(CMX_E_ARG, "trace too large: raise trace_every");
        if ((rc = B.upload(&d_trace, std::vector<double>(n_samples * (size_t)n, std::nan(""))))) return rc;
        a.trace = d_trace;
        a.trace_every = c.trace_every;
        a.trace_n = n;
    }

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
    const bool rl = c.integrator == CMX_RUSH_LARSEN;
    auto launch = [&](int k0, int k1, int lanes) -> int {
        // only the cells still running: the sorted prefix with n_steps > k0
        // (counts sorted descending: lower_bound with greater<> finds the first <= k0)
        CellsArgs<RS> ak = a;
        ak.n = (int)(std::lower_bound(s_steps.begin(), s_steps.end(), k0, std::greater<int32_t>()) - s_steps.begin());
        if (ak.n <= 0) return 0;
        batch_launch<RS>(model_ops(c.model))(ak, rl, k0, k1, stream, lanes);
        return 1;
    };
    // lanes_per_cell 0: per launch, two wavefronts per 64 cells while the
    // launch's one-lane wavefronts would fill at most half of the device's
    // SIMDs (each then hosts one wavefront of a cell pair, and a cell's step
    // takes the longer of its two halves); one while more run (a SIMD then
    // interleaves wavefronts anyway and the split only adds the exchange)
    const int lanes_cfg = c.lookup_tables ? 1 : c.lanes_per_cell;
    int n_simd = 1024;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            n_simd = 4 * prop.multiProcessorCount;
        else
            (void)hipGetLastError();
    }
    int split_launches = 0;
    CELLS_HIP(hipEventRecord(e0, stream));
    for (int k0 = 0; k0 < longest; k0 += chunk) {
        const int active = (int)(std::lower_bound(s_steps.begin(), s_steps.end(), k0, std::greater<int32_t>()) -
                                 s_steps.begin());
        const int waves = (active + 63) / 64;
        const int lanes = lanes_cfg == 0 ? (2 * waves <= n_simd ? 2 : 1) : lanes_cfg;
        if (!launch(k0, std::min(longest, k0 + chunk), lanes)) break;
        CELLS_HIP(hipGetLastError());
        ++launches;
        split_launches += lanes == 2;
    }
    CELLS_HIP(hipEventRecord(e1, stream));
    CELLS_HIP(hipStreamSynchronize(stream));
    float ms = 0.f;
    CELLS_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (d_miss) {
        int32_t miss = 0;
        CELLS_HIP(hipMemcpy(&miss, d_miss, sizeof(miss), hipMemcpyDeviceToHost));
        if (miss)
            return cfail(CMX_E_NONFINITE, "cell " + std::to_string(order[miss - 1]) +
                                              ": membrane voltage outside the gate lookup table (-250..350 mV)");
    }

    // back to the caller's order
    CELLS_HIP(hipMemcpy(s_state.data(), d_state, s_state.size() * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(s_v.data(), d_v, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> cnt(ps);
    std::vector<double> on(ps), lp(ps), en(ps);
    CELLS_HIP(hipMemcpy(cnt.data(), a.ap_count, ps * sizeof(int32_t), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(on.data(), a.onset, ps * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(lp.data(), a.last_peak, ps * sizeof(double), hipMemcpyDeviceToHost));
    CELLS_HIP(hipMemcpy(en.data(), a.end, ps * sizeof(double), hipMemcpyDeviceToHost));
    if (d_trace) {
        std::vector<double> tr(n_samples * (size_t)n);
        CELLS_HIP(hipMemcpy(tr.data(), d_trace, tr.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t s = 0; s < n_samples; ++s)
            for (int j = 0; j < n; ++j) c.trace[s * (size_t)n + order[j]] = tr[s * (size_t)n + j];
    }
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
        stats->split_launches = split_launches;
    }
    return CMX_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free variant (cmx_cells_run_noalloc): every buffer the
// kernels work in -- state and parameter planes, step counts, padded
// stimuli, the CellProperties accumulators -- is carved out of one static
// device array instead of hipMalloc'd storage.  Its device address is looked
// up once per device (hipGetSymbolAddress); host code never passes the
// array's name itself to a kernel.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = CMX_CELLS_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on the current device (looked up on the
// device's first call).
int scratch_ptr(void** out)
{
    int dev = 0;
    CELLS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return cfail(CMX_E_ARG, "device index beyond the supported range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CELLS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return cfail(CMX_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CMX_OK;
}

// Layout of n cells in the scratch array: 256-byte aligned regions of
// ps = n rounded up to 64 elements each.
struct ScratchLayout {
    size_t state, V, params, pdef, steps, stim, phase, count, onset, peak, last_peak, end, bytes;
};

ScratchLayout scratch_layout(int n, int S, int P, size_t rs)
{
    const size_t ps = ((size_t)n + 63) / 64 * 64;
    ScratchLayout L{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t off = at;
        at += (bytes + 255) / 256 * 256;
        return off;
    };
    L.state = take((size_t)(S - 1) * ps * sizeof(double));
    L.V = take(ps * sizeof(double));
    L.params = take((size_t)P * ps * rs);
    L.pdef = take((size_t)std::max(P, 1) * rs);
    L.steps = take(ps * sizeof(int32_t));
    L.stim = take(ps * sizeof(cmx_stimulus));
    L.phase = take(ps * sizeof(int32_t));
    L.count = take(ps * sizeof(int32_t));
    L.onset = take(ps * sizeof(double));
    L.peak = take(ps * sizeof(double));
    L.last_peak = take(ps * sizeof(double));
    L.end = take(ps * sizeof(double));
    L.bytes = at;
    return L;
}

struct ModelConsts {
    double init[kMaxModelStates];
    double pdef[kMaxModelStates];
};

// One thread per cell: step count from t_end (0, and the cell marked invalid
// with -1, when t_end is not a whole number of dt), initial state (y0 or the
// model's), parameters, padded stimulus, and every CellProperties
// accumulator reset -- nothing is left from an earlier call.
template <class RS>
__global__ __launch_bounds__(kBlock) void noalloc_prepare_kernel(CellsArgs<RS> a, int S, int P, const double* t_end,
                                                                 const cmx_stimulus* stim_in, const double* params_in,
                                                                 const double* y0, ModelConsts mc, RS* pdef,
                                                                 int32_t* steps, cmx_stimulus* stim)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < P) pdef[i] = (RS)mc.pdef[i];
    if (i >= a.n) return;
    const int n = a.n;
    const double r = t_end[i] / a.dt;
    const double k = rint(r);
    bool ok = t_end[i] >= 0.0 && fabs(r - k) <= 1e-9 * fmax(1.0, r) && k <= 2147483647.0;
    if (stim_in) {
        cmx_stimulus s = stim_in[i];
        if (s.period > 0) {
            ok = ok && s.duration <= s.period;
            s.duration += s.period * DBL_EPSILON;
        } else {
            s.duration += (s.duration + s.start) * DBL_EPSILON;
        }
        if (!(s.stop > 0)) s.stop = DBL_MAX;
        stim[i] = s;
    }
    steps[i] = ok ? (int32_t)k : 0;
    a.V[i] = y0 ? y0[i] : mc.init[0];
    for (int s = 1; s < S; ++s) a.state[(size_t)(s - 1) * a.ps + i] = y0 ? y0[(size_t)s * n + i] : mc.init[s];
    for (int q = 0; q < a.n_param_planes; ++q) ((RS*)a.params)[(size_t)q * a.ps + i] = (RS)params_in[(size_t)q * n + i];
    a.phase[i] = ok ? 0 : -1;
    a.ap_count[i] = 0;
    a.onset[i] = -1.0;
    a.peak[i] = -DBL_MAX;
    a.last_peak[i] = -1.0;
    a.end[i] = -1.0;
}

// Final states [S][n] and the summary [4][n] in the caller's buffers; an
// invalid cell (phase -1 before the run, left untouched by it since it takes
// no step) gets NaN everywhere.
__global__ __launch_bounds__(kBlock) void noalloc_finish_kernel(CellsArgs<double> a, int S, double* y_out, double* props)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int n = a.n;
    const bool bad = a.phase[i] < 0;
    const double nan = __builtin_nan("");
    if (y_out) {
        y_out[i] = bad ? nan : a.V[i];
        for (int s = 1; s < S; ++s) y_out[(size_t)s * n + i] = bad ? nan : a.state[(size_t)(s - 1) * a.ps + i];
    }
    if (props) {
        props[i] = bad ? nan : (double)a.ap_count[i];
        props[(size_t)n + i] = bad ? nan : a.onset[i];
        props[2 * (size_t)n + i] = bad ? nan : a.last_peak[i];
        props[3 * (size_t)n + i] = bad ? nan : a.end[i];
    }
}

template <class RS>
int run_cells_noalloc(const cmx_cells& c, int S, int P, const double* y0, double* y_out, double* props, hipStream_t stream)
{
    const int n = c.n_cells;
    const ScratchLayout L = scratch_layout(n, S, P, sizeof(RS));
    if (L.bytes > kScratchBytes) {
        int cap = 0;
        for (int step = 1 << 30; step > 0; step >>= 1)
            if (scratch_layout(cap + step, S, P, sizeof(RS)).bytes <= kScratchBytes) cap += step;
        return cfail(CMX_E_ARG, "n_cells = " + std::to_string(n) + " exceeds the static scratch array (" +
                                    std::to_string(kScratchBytes) + " bytes holds at most " + std::to_string(cap) +
                                    " cells of this model and precision)");
    }
    void* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    char* b = (char*)base;
    CellsArgs<RS> a{};
    a.n = n;
    a.ps = (n + 63) / 64 * 64;
    a.dt = c.dt;
    a.cm = 1.0;
    a.thr = c.threshold;
    a.n_steps = (const int32_t*)(b + L.steps);
    a.stim = c.stimuli ? (const cmx_stimulus*)(b + L.stim) : nullptr;
    a.state = (double*)(b + L.state);
    a.V = (double*)(b + L.V);
    a.params = c.n_params > 0 ? (const RS*)(b + L.params) : nullptr;
    a.pdef = (const RS*)(b + L.pdef);
    a.n_param_planes = c.n_params;
    a.phase = (int32_t*)(b + L.phase);
    a.ap_count = (int32_t*)(b + L.count);
    a.onset = (double*)(b + L.onset);
    a.peak = (double*)(b + L.peak);
    a.last_peak = (double*)(b + L.last_peak);
    a.end = (double*)(b + L.end);
    ModelConsts mc{};
    ModelMeta mm{};
    model_meta(c.model, &mm);
    const double* init = mm.init;
    const double* pd = mm.pdef;
    for (int s = 0; s < S; ++s) mc.init[s] = init[s];
    for (int q = 0; q < P; ++q) mc.pdef[q] = pd[q];
    const int grid = (n + kBlock - 1) / kBlock;
    noalloc_prepare_kernel<RS><<<grid, kBlock, 0, stream>>>(a, S, P, c.t_end, c.stimuli, c.params, y0, mc,
                                                            (RS*)(b + L.pdef), (int32_t*)(b + L.steps),
                                                            (cmx_stimulus*)(b + L.stim));
    CELLS_HIP(hipGetLastError());
    // every cell runs its own step count in one launch (k1 beyond any count)
    const bool rl = c.integrator == CMX_RUSH_LARSEN;
    const int lanes = c.lanes_per_cell == 2 ? 2 : 1;
    batch_launch<RS>(model_ops(c.model))(a, rl, 0, INT32_MAX, stream, lanes);
    CELLS_HIP(hipGetLastError());
    CellsArgs<double> f{};
    f.n = a.n;
    f.ps = a.ps;
    f.state = a.state;
    f.V = a.V;
    f.phase = a.phase;
    f.ap_count = a.ap_count;
    f.onset = a.onset;
    f.last_peak = a.last_peak;
    f.end = a.end;
    noalloc_finish_kernel<<<grid, kBlock, 0, stream>>>(f, S, y_out, props);
    CELLS_HIP(hipGetLastError());
    return CMX_OK;
}

}  // namespace
}  // namespace cmx

using namespace cmx;

extern "C" {

int cmx_cells_run(const cmx_cells* c, const double* y0, double* y_out, double* props, cmx_cells_stats* stats)
{
    if (!c) return cfail(CMX_E_ARG, "null description");
    ModelMeta mm{};
    if (!model_meta(c->model, &mm)) return cfail(CMX_E_ARG, "unknown cell model");
    const int S = mm.S, P = mm.P;
    if (c->integrator != CMX_FORWARD_EULER && c->integrator != CMX_RUSH_LARSEN)
        return cfail(CMX_E_ARG, "batched cells support forward-Euler and Rush-Larsen integration");
    if (c->precision != CMX_FP64 && c->precision != CMX_MIXED) return cfail(CMX_E_ARG, "unknown precision");
    if (c->n_cells < 0 || (c->n_cells > 0 && !c->t_end)) return cfail(CMX_E_ARG, "n_cells and t_end are required");
    if (!(c->dt > 0)) return cfail(CMX_E_ARG, "dt must be positive");
    if (c->n_params < 0 || c->n_params > P || (c->n_params > 0 && !c->params))
        return cfail(CMX_E_ARG, "n_params must be in [0, the model's parameter count] with params given");
    if (c->lanes_per_cell < 0 || c->lanes_per_cell > 2) return cfail(CMX_E_ARG, "lanes_per_cell must be 0, 1 or 2");
    if (c->lookup_tables && c->lanes_per_cell == 2) return cfail(CMX_E_ARG, "lookup tables need lanes_per_cell 1");
    if (c->trace_every < 0 || (c->trace_every > 0 && !c->trace))
        return cfail(CMX_E_ARG, "trace_every must be >= 0, with a trace array when > 0");
    if (c->n_cells == 0) return CMX_OK;
    return c->precision == CMX_MIXED ? run_cells<float>(*c, S, P, y0, y_out, props, stats)
                                     : run_cells<double>(*c, S, P, y0, y_out, props, stats);
}

int cmx_cells_run_noalloc(const cmx_cells* c, const double* y0, double* y_out, double* props, void* stream)
{
    if (!c) return cfail(CMX_E_ARG, "null description");
    ModelMeta mm{};
    if (!model_meta(c->model, &mm)) return cfail(CMX_E_ARG, "unknown cell model");
    const int S = mm.S, P = mm.P;
    if (S > kMaxModelStates || P > kMaxModelStates) return cfail(CMX_E_ARG, "model too large for the variant");
    if (c->integrator != CMX_FORWARD_EULER && c->integrator != CMX_RUSH_LARSEN)
        return cfail(CMX_E_ARG, "batched cells support forward-Euler and Rush-Larsen integration");
    if (c->precision != CMX_FP64 && c->precision != CMX_MIXED) return cfail(CMX_E_ARG, "unknown precision");
    if (c->n_cells < 0 || (c->n_cells > 0 && !c->t_end)) return cfail(CMX_E_ARG, "n_cells and t_end are required");
    if (!(c->dt > 0)) return cfail(CMX_E_ARG, "dt must be positive");
    if (c->n_params < 0 || c->n_params > P || (c->n_params > 0 && !c->params))
        return cfail(CMX_E_ARG, "n_params must be in [0, the model's parameter count] with params given");
    if (c->lookup_tables) return cfail(CMX_E_ARG, "lookup tables need a table allocation: use cmx_cells_run");
    if (c->trace_every != 0) return cfail(CMX_E_ARG, "a voltage trace needs an allocation: use cmx_cells_run");
    if (c->lanes_per_cell < 0 || c->lanes_per_cell > 2) return cfail(CMX_E_ARG, "lanes_per_cell must be 0, 1 or 2");
    if (c->n_cells == 0) return CMX_OK;
    hipStream_t s = (hipStream_t)stream;
    return c->precision == CMX_MIXED ? run_cells_noalloc<float>(*c, S, P, y0, y_out, props, s)
                                     : run_cells_noalloc<double>(*c, S, P, y0, y_out, props, s);
}

const char* cmx_cells_last_error(void) { return g_cells_error.c_str(); }

}  // extern "C"
