// The context behind the C ABI's opaque cmx_ctx and the cell-model wrappers,
// shared by the time-step runtime (cmx_device.hip) and the adjoint
// (cmx_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "cmx_internal.h"
#include "cmx_kernels.h"
#include "models/dn85.h"
#include "models/fox2002.h"
#include "models/fr2000.h"
#include "models/hh52.h"
#include "models/lr91.h"
#include "models/mahajan2008.h"
#include "models/maleckar2008.h"
#include "models/n98.h"
#include "models/ord_endo.h"
#include "models/tt06_epi.h"

namespace cmx {

// ---------------------------------------------------------------------------
// Cell models
// ---------------------------------------------------------------------------
// A kernel-side model: the generated (or hand-written) functions of namespace
// ns behind one interface.  NS states (state 0 is V), NP parameters; eval
// gives the right-hand side, and with RL the Rush-Larsen steady states and
// rates of the gate states (gate(i)); be_step / grl_step (generated models
// only) are the backward-Euler and generalized Rush-Larsen substeps.
#define CMX_MODEL_COMMON(ns)                                                                                      \
    static constexpr int NS = ns::kNumStates;                                                                     \
    static constexpr int NP = ns::kNumParams;                                                                     \
    template <class R, bool RL>                                                                                   \
    static CMX_HD __forceinline__ void eval(const R* y, R V, R st, const R* p, R cm, R* dy, R* gi, R* gk)         \
    {                                                                                                             \
        ns::evaluate<R, RL>(y, V, st, p, cm, dy, gi, gk);                                                         \
    }                                                                                                             \
    template <class R>                                                                                            \
    static __device__ __forceinline__ R iion(const R* y, R V, const R* p, R cm) { return ns::iionic<R>(y, V, p, cm); } \
    template <class R>                                                                                            \
    static __device__ __forceinline__ R dvdt(const R* y, R V, R st, const R* p, R cm)                             \
    {                                                                                                             \
        return ns::dvdt<R>(y, V, st, p, cm);                                                                      \
    }                                                                                                             \
    static constexpr bool gate(int i) { return ns::kIsGate[i] != 0; }                                            \
    static const double* initial_state() { return ns::kInitialState; }                                           \
    static const double* param_defaults() { return ns::kParamDefaults; }                                         \
    static const char* const* state_names() { return ns::kStateNames; }                                          \
    static const char* const* param_names() { return ns::kParamNames; }                                          \
    static const char* const* param_aliases() { return ns::kParamAliases; }

#define CMX_MODEL_IMPLICIT(ns)                                                                                    \
    template <class R>                                                                                            \
    static __device__ __forceinline__ int be_step(R* y, R V, R st, const R* p, R cm, R dt)                        \
    {                                                                                                             \
        return ns::be_step<R>(y, V, st, p, cm, dt);                                                               \
    }                                                                                                             \
    template <class R, int ORDER>                                                                                 \
    static __device__ __forceinline__ void grl_step(R* y, R st, const R* p, R cm, R dt, bool update_v)            \
    {                                                                                                             \
        ns::grl_step<R, ORDER>(y, st, p, cm, dt, update_v);                                                       \
    }

// A generated model (all four integrators).
#define CMX_GENERATED_MODEL(Name, ns) \
    struct Name {                     \
        CMX_MODEL_COMMON(ns)          \
        CMX_MODEL_IMPLICIT(ns)        \
    };

CMX_GENERATED_MODEL(ModelLr91, lr91)
CMX_GENERATED_MODEL(ModelTt06, tt06_epi)
CMX_GENERATED_MODEL(ModelHh52, hh52)
CMX_GENERATED_MODEL(ModelFox2002, fox2002)
CMX_GENERATED_MODEL(ModelMahajan2008, mahajan2008)
CMX_GENERATED_MODEL(ModelDn85, dn85)
CMX_GENERATED_MODEL(ModelFr2000, fr2000)
CMX_GENERATED_MODEL(ModelMaleckar2008, maleckar2008)
CMX_GENERATED_MODEL(ModelN98, n98)

// O'Hara-Rudy 2011 endocardial cells: forward Euler and Rush-Larsen only (no
// backward-Euler or generalized Rush-Larsen forms).
struct ModelOrd {
    CMX_MODEL_COMMON(ord_endo)
};

// Gate columns of a model's lookup table: gate state s (Model::gate(s)) has
// its steady state in column 2 col(s) and its decay factor in 2 col(s) + 1;
// rows are padded to whole float4s.
template <class Model>
struct GateCols {
    static constexpr int count()
    {
        int c = 0;
        for (int i = 1; i < Model::NS; ++i) c += Model::gate(i) ? 1 : 0;
        return c;
    }
    static constexpr int col(int s)
    {
        int c = 0;
        for (int i = 1; i < s; ++i) c += Model::gate(i) ? 1 : 0;
        return c;
    }
    static constexpr int width() { return (2 * count() + 3) / 4 * 4; }
};

// Rush-Larsen coefficients of every gate at voltage V (host, FP64): the
// model's right-hand side at its initial state with parameters p.
template <class Model>
void gate_coefficients(double V, const double* p, double cm, double* gi, double* gk)
{
    double dy[64] = {0.0};
    Model::template eval<double, true>(Model::initial_state(), V, 0.0, p, cm, dy, gi, gk);
}

// Gate lookup tables (mixed-precision Rush-Larsen cells): for
// V = kLutVmin + r kLutDv mV, each gate's steady state and its decay factor
// exp(g_k dt) over one ODE step of length dt, evaluated in FP64 by the model
// itself and stored as floats in the GateCols layout.  The tables hold
// V-only quantities, so a model whose gate coefficients change with its
// parameters is refused.
constexpr double kLutVmin = -250.0, kLutVmax = 350.0, kLutDv = 0.02;

template <class Model>
const char* fill_gate_table(double dt, double cm, std::vector<float>& table, int* rows_out)
{
    const double* pdef = Model::param_defaults();
    constexpr int P = Model::NP;
    constexpr int W = GateCols<Model>::width();
    const int rows = (int)((kLutVmax - kLutVmin) / kLutDv + 0.5) + 1;
    table.assign((size_t)rows * W, 0.0f);
    std::vector<double> p(pdef, pdef + P), p2(P + 1);
    for (int q = 0; q < P; ++q) p2[q] = 2.0 * pdef[q] + 1.0;
    double gi[64], gk[64], gi2[64], gk2[64];
    for (int r = 0; r < rows; ++r) {
        const double V = kLutVmin + r * kLutDv;
        gate_coefficients<Model>(V, p.data(), cm, gi, gk);
        if (r % 500 == 0) {
            gate_coefficients<Model>(V, p2.data(), cm, gi2, gk2);
            for (int s = 1; s < Model::NS; ++s)
                if (Model::gate(s) && (gi2[s] != gi[s] || gk2[s] != gk[s]))
                    return "lookup tables: this model's gate coefficients depend on its parameters";
        }
        for (int s = 1; s < Model::NS; ++s) {
            if (!Model::gate(s)) continue;
            const int col = 2 * GateCols<Model>::col(s);
            table[(size_t)r * W + col] = (float)gi[s];
            table[(size_t)r * W + col + 1] = (float)std::exp(gk[s] * dt);
            if (!std::isfinite(table[(size_t)r * W + col]) || !std::isfinite(table[(size_t)r * W + col + 1]))
                return "lookup tables: a gate coefficient is not finite inside the table's voltage range";
        }
    }
    *rows_out = rows;
    return nullptr;
}

// Every gate's Rush-Larsen steady state and decay factor at V from a table
// (rows at vmin + r / inv_dv), linearly interpolated between the two rows
// around V.  Returns false when V lies outside the table (the values then
// come from its first rows).
template <class Model>
__device__ __forceinline__ bool lut_gates(const float* __restrict__ lut, float vmin, float inv_dv, int rows, float V,
                                          float* ginf, float* gdec)
{
    constexpr int W = GateCols<Model>::width();
    const float x = (V - vmin) * inv_dv;
    const bool inside = x >= 0.0f && x <= (float)(rows - 1);
    const float xs = inside ? x : 0.0f;
    int r = (int)xs;
    r = r > rows - 2 ? rows - 2 : r;
    const float f = xs - (float)r;
    const float4* t0 = reinterpret_cast<const float4*>(lut + (size_t)r * W);
    const float4* t1 = reinterpret_cast<const float4*>(lut + (size_t)(r + 1) * W);
    float lo[W], hi[W];
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
        const float4 u = t0[q], v = t1[q];
        lo[4 * q] = u.x; lo[4 * q + 1] = u.y; lo[4 * q + 2] = u.z; lo[4 * q + 3] = u.w;
        hi[4 * q] = v.x; hi[4 * q + 1] = v.y; hi[4 * q + 2] = v.z; hi[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int s = 1; s < Model::NS; ++s) {
        if (!Model::gate(s)) continue;
        const int c = 2 * GateCols<Model>::col(s);
        ginf[s] = lo[c] + f * (hi[c] - lo[c]);
        gdec[s] = lo[c + 1] + f * (hi[c + 1] - lo[c + 1]);
    }
    return inside;
}

}  // namespace cmx

// Cells of one (model, integrator) pair when a tissue mixes cell types
// (AbstractCardiacCellFactory::CreateCardiacCellForTissueNode returning
// different classes per node): cell j of the group sits at node h_node[j]
// (ascending), its state planes are indexed by j.
struct CellGroup {
    int model = 0, integrator = 0, S = 0, P = 0, n = 0, ps = 0;
    int32_t* node = nullptr;     // device copy of h_node
    void* state = nullptr;       // [S-1][ps] RS
    void* pdef = nullptr;        // [P] RS
    std::vector<int32_t> h_node;
};

struct cmx_ctx {
    std::string err;
    int n = 0, dim = 0, model = 0, integrator = 0, precision = 0, format = 0;
    int S = 0, P = 0, n_param_planes = 0, param_planes_alloc = 0;
    double dt = 0, dt_ode = 0, am = 0, cm = 0, rtol = 0, atol = 0;
    int max_its = 0, refine = 0, n_sub = 1, use_graph = 0, extrapolate = 0;
    int split = 0, n_half = 1;   // operator splitting: ODE substeps per half step
    int width = 0, n_patterns = 0;
    int G = 1;                   // blocks of the reduction kernels
    int cg_rows = 0;             // CG kernels' rows per thread held in registers (0: precision default)
    int graph_steps = 8;         // steps in the largest captured graph (a power of two)
    // adjoint: bound on its stored forward states (bytes, 0 = free memory) and
    // the plan of its last call
    double adj_mem_limit = 0.0;
    int adj_slots = 0, adj_sub = 0;
    long long adj_iterations = 0;  // CG iterations of the adjoint's solves (all calls)
    // gate lookup table (mixed Rush-Larsen cells): rows at lut_vmin + r lut_dv
    float* lut = nullptr;
    int lut_rows = 0;
    double lut_vmin = 0.0, lut_dv = 0.0;
    long long nnz = 0;
    hipStream_t stream = nullptr;
    // device state
    double* V = nullptr;
    void* state = nullptr;       // [S-1][n] RS
    void* params = nullptr;      // [P][n] RS
    void* pdef = nullptr;        // [P] RS
    // more than one cell type: one group per (model, integrator); state,
    // params and pdef above are then unused (model / integrator / S / P are
    // the first group's)
    std::vector<CellGroup> cell_groups;
    uint8_t* stim_id = nullptr;
    cmx_stimulus* stim = nullptr;
    void* F = nullptr;
    void* vec[15] = {};          // r w z s p m_a m_b x1 x2 x0 x_{n-2} r2 p2 xf u
    double* Y = nullptr;         // V^n + x0 (initial guess)
    void* G1 = nullptr;          // F - c x0
    double* r0_64 = nullptr;     // FP64 step residual (refinement)
    double* iion = nullptr;      // scratch for cmx_get_iionic
    // state-variable interpolation
    int svi = 0;
    void* svi_iion = nullptr;    // nodal Iion of the step (RV)
    void* svi_out = nullptr;     // correction of b (RV)
    int32_t* elems = nullptr;
    double* edet = nullptr;
    int32_t* inc_ptr = nullptr;
    int32_t* inc = nullptr;
    double qw[4] = {}, qp[4][3] = {};
    int nq = 0;
    double* act = nullptr;       // [5][n] activation times (cmx_set_activation)
    // the adjoint's replay: ICI steps [trace_k0, trace_k0 + trace_len) write the
    // state they start from to trace (slots of trace_slot doubles); plain launches only
    double* trace = nullptr;
    int trace_k0 = 0, trace_len = 0;
    size_t trace_slot = 0;
    double act_thr = 0.0;
    int act_every = 0;
    // matrix
    uint8_t* pid = nullptr;
    std::vector<int32_t> h_off;  // stencil offsets, padded to the kernel width
    int pad = 0;                 // guard elements on each side of the gathered vectors
    int ps = 0;                  // cell-state / parameter plane stride (elements)
    int32_t* ell_col = nullptr;
    uint8_t* ell_set = nullptr;      // offset-set ELL: set of each row
    int32_t* ell_off = nullptr;      // [sets][width] offsets
    double *m64 = nullptr, *k64 = nullptr, *a64 = nullptr, *dinv64 = nullptr;
    float *a32 = nullptr, *dinv32 = nullptr;
    double* partials = nullptr;  // 3 buffers of [kMaxBlocks * 3]
    cmx::StepClock* clock = nullptr;
    cmx::SolveCtl* ctl = nullptr;     // [2]: first solve, refinement solve
    double2* hist = nullptr;
    int32_t* its_log = nullptr;  // [batch][4] per-solve iterations, indexed by step % batch
    // pinned host mirrors
    cmx::StepClock* h_clock = nullptr;
    cmx::SolveCtl* h_ctl = nullptr;
    int32_t* h_its = nullptr;
    int batch = 64;
    // host copies for cmx_get_iionic
    std::vector<cmx_stimulus> h_stim;
    std::vector<uint8_t> h_stim_id;
    double last_t_next = 0.0;
    bool stepped = false;
    // CG iteration budgets (first solve, refinement solve) per enqueued step
    int budget = 8;
    int refine_budget = 1;
    bool fold_refine = true;     // the refinement residual in the last iteration when it can be
    // stats
    long long steps = 0, cg_total = 0, refine_total = 0, resumes = 0, folded_total = 0;
    int last_its = 0;
    double last_residual = 0;
    double device_bytes = 0;
    std::vector<void*> allocs;
    // hipGraphs of whole steps, keyed by (budget, refinement budget, steps,
    // solver blocks, CG rows per thread, activation tracking)
    struct StepGraph {
        int budget, refine_budget, steps, G, rows, act_every;
        double act_thr;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    long long graph_launches = 0, graph_captures = 0, graph_evictions = 0;
    int graph_profiler_off = 0;  // a profiler's dispatch interception was detected: plain launches
    // kernel timing (cmx_set_profiling)
    int prof_mask = 0;
    struct Pending { int kind; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double k_ms[CMX_K_COUNT] = {};
    long long k_launches[CMX_K_COUNT] = {};
};

namespace cmx {

// The clock of one run of n_total steps from t0 (TimeStepper(t0, t0 + n_total dt, dt)).
inline StepClock run_clock(double t0, double dt, int n_total)
{
    StepClock ck{};
    ck.t0 = t0;
    ck.dt = dt;
    ck.t_end = t0 + n_total * dt;
    ck.eps = ck.t_end > 1.0 ? DBL_EPSILON * ck.t_end : DBL_EPSILON;
    ck.k = 0;
    ck.n_steps = n_total;
    return ck;
}

// SimpleStimulus pads its duration by (duration + start) eps and
// RegularStimulus by period eps (heart/src/stimulus/SimpleStimulus.cpp:44-52,
// RegularStimulus.cpp:40-56); a stop time <= 0 means none.  False when a
// regular stimulus lasts longer than its period.
inline bool pad_stimulus(cmx_stimulus& s)
{
    if (s.period > 0) {
        if (s.duration > s.period) return false;
        s.duration += s.period * DBL_EPSILON;
    } else {
        s.duration += (s.duration + s.start) * DBL_EPSILON;
    }
    if (!(s.stop > 0)) s.stop = DBL_MAX;
    return true;
}

// Steps k0 .. k1-1 of the run that cmx_step(c, t0, n_total) takes, with the
// same step times (the adjoint replays a run step by step).
int step_range(cmx_ctx* c, double t0, int n_total, int k0, int k1);

// A x = b from x = 0 with the step's pipelined Jacobi-PCG kernels, on the
// caller's buffers (FP64 contexts; the adjoint's solves).  Stopping test
// ||D^-1 r|| <= rtol ||D^-1 b||.  plain_cg_start launches the start, which
// ends a solve whose b is already below the test; plain_cg_iters launches
// iterations from .. from+count-1, each returning at once once ctl[0].done
// is set (1 converged, 2 max_its reached, 3 non-finite).
// The ELL rows of a context (column indices, or offset sets).
inline EllRows ell_rows(const cmx_ctx* c)
{
    EllRows R{c->ell_col, c->width, c->n};
    R.sets = c->format == CMX_FMT_ELL_OFFSETS ? c->ell_set : nullptr;
    R.set_off = c->ell_off;
    return R;
}

struct PlainCg {
    double *r, *w, *z, *s;
    double* x;           // the solution; gathered (guard zones) when a solve starts from it
    int warm;            // start from the x the vector holds (r = b - A x) instead of x = 0
    double* p;           // gathered at the start: c->pad zeroed guard elements on each side
    double* m[2];        // gathered, likewise
    double* part[3];     // [kMaxBlocks * 3] each: iteration ping-pong, start
    double2* hist;       // [max_its + 1]
    SolveCtl* ctl;       // [2]
    double rtol;
    int max_its;
};
int plain_cg_start(cmx_ctx* c, const PlainCg& s, const double* b);
int plain_cg_iters(cmx_ctx* c, const PlainCg& s, int from, int count);

}  // namespace cmx
