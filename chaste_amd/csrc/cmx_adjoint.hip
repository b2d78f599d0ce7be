// Discrete adjoint of the FP64 monodomain time-step (ICI form, forward-Euler
// or Rush-Larsen cells) behind cmx_adjoint.  The reference has no adjoint;
// the forward map differentiated here is exactly the step the runtime takes
// (cmx_device.hip), restated per node as
//   y^{n+1} = Phi(y^n, V^n, p)          cell substeps from t_n to t_{n+1}, V clamped
//   I^n     = Iion(y^{n+1}, V^n, p)     (AbstractCardiacTissue::SolveCellSystems)
//   A V^{n+1} = M (c V^n - Am I^n - Istim(t_{n+1})),  c = Am Cm / dt,  A = c M + K
// (MonodomainSolver::SetupLinearSystem, LinearSystem::Solve).  For the loss
// L = sum over observation steps j and nodes o of (V^j_o - d^j_o)^2, with
// lambda_V^n = dL/dV^n and lambda_y^n = dL/dy^n through the later steps:
//   mu       = A^{-1} lambda_V^{n+1}     (A is symmetric: the transposed solve is a solve)
//   l_z      = M mu
//   lambda_y^n = J_y^T lambda_y^{n+1} - Am l_z dI/dy^n
//   lambda_V^n = J_V^T lambda_y^{n+1} + l_z (c - Am dI/dV^n) + 2 (V^n - d^n) at observed nodes
//   dL/dp   += J_p^T lambda_y^{n+1} - Am l_z dI/dp
// The Jacobian columns of (Phi, Iion) come from forward-mode differentiation
// of the generated cell model (four seeded inputs per pass, cmx_dualn); the
// states of each step are recomputed from checkpoints.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cmx_models.h"

namespace cmx {
namespace {

constexpr int kAdjChunk = 16;       // CG iterations of a solve's first launch batch (no history yet)
constexpr int kAdjFirstMax = 64;    // cap on the first batch taken from the previous solve's count
constexpr int kAdjMaxIts = 2000;

// lambda_V at the last step: the observation term only; lambda_y = 0.
__global__ void adjoint_final_kernel(int n, int ps, int ns1, const double* __restrict__ V,
                                     const int32_t* __restrict__ obs_slot, const double* __restrict__ obs_d,
                                     double* __restrict__ lam_v, double* __restrict__ lam_y)
{
    const int i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    double lv = 0.0;
    if (obs_d && obs_slot[i] >= 0) lv = 2.0 * (V[i] - obs_d[obs_slot[i]]);
    lam_v[i] = lv;
    for (int s = 0; s < ns1; ++s) lam_y[(size_t)s * ps + i] = 0.0;
}

__global__ void gather_kernel(int m, const int32_t* __restrict__ idx, const double* __restrict__ V,
                              double* __restrict__ out)
{
    const int o = blockIdx.x * kAdjBlock + threadIdx.x;
    if (o < m) out[o] = V[idx[o]];
}

// ---------------------------------------------------------------------------
// The adjoint's linear solves A mu = b (A symmetric) run on the forward
// step's pipelined Jacobi-PCG kernels (cmx::plain_cg_start / plain_cg_iters,
// one kernel per iteration); l_z = M mu is one more product.
// ---------------------------------------------------------------------------
template <class Rows>
__global__ __launch_bounds__(kCgBlock) void adj_spmv_kernel(Rows R, const double* __restrict__ vals, int n,
                                                            const double* __restrict__ x, double* __restrict__ y)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const RowRange rr(n, blk, gridDim.x);
    for (int i = rr.lo + threadIdx.x; i < rr.hi; i += kCgBlock) {
        const typename Rows::Row row = R.row(i);
        double s = 0.0;
        for (int k = 0; k < R.width(); ++k) s += vals[R.vidx(row, k)] * x[R.col(row, k)];
        y[i] = s;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int afail(cmx_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

#define ADJ_HIP(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return afail(ctx, CMX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// Device buffers of one cmx_adjoint call, freed on every return path.
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(cmx_ctx* c, T** out, size_t count)
    {
        void* p = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return afail(c, CMX_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed (adjoint)");
        }
        ptrs.push_back(p);
        *out = (T*)p;
        return CMX_OK;
    }
    // A vector the stencil gathers: c->pad zeroed guard elements on each side.
    int alloc_gathered(cmx_ctx* c, double** out, size_t count)
    {
        double* base = nullptr;
        const size_t total = count + 2 * (size_t)c->pad;
        int rc = alloc(c, &base, total);
        if (rc) return rc;
        if (hipMemset(base, 0, total * sizeof(double)) != hipSuccess)
            return afail(c, CMX_E_HIP, "hipMemset of an adjoint vector failed");
        *out = base + c->pad;
        return CMX_OK;
    }
};

template <int W>
StencilRows<W> adj_stencil_rows(cmx_ctx* c)
{
    StencilRows<W> R;
    R.pid = c->pid;
    R.n = c->n;
    R.n_patterns = c->n_patterns;
    R.pad = c->pad;
    for (int k = 0; k < W; ++k) R.off[k] = c->h_off[k];
    return R;
}

struct AdjBuffers {
    double *lam_v, *lam_y, *lz, *grad;
    const double* b;             // right-hand side of the solve: lambda_V^{n+1}
    PlainCg cg;                  // the solve's vectors, reductions and control
    SolveCtl* h_ctl;             // pinned mirror of cg.ctl[0]
    int last_its;                // iterations of the previous solve
};

// mu = A^{-1} b (b = lambda_V^{n+1}) into B.cg.x, then l_z = M mu.
template <class Rows>
int adjoint_solve(cmx_ctx* c, const Rows& R, AdjBuffers& B)
{
    int rc = plain_cg_start(c, B.cg, B.b);
    if (rc) return rc;
    // launch as many iterations as the previous solve took (plus the launch
    // that detects convergence), then a few at a time
    int it = 0, chunk = B.last_its > 0 ? std::min(B.last_its + 1, kAdjFirstMax) : kAdjChunk;
    for (;;) {
        if ((rc = plain_cg_iters(c, B.cg, it, chunk))) return rc;
        it += chunk;
        ADJ_HIP(c, hipMemcpyAsync(B.h_ctl, B.cg.ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
        ADJ_HIP(c, hipStreamSynchronize(c->stream));
        if (B.h_ctl->done == 1) {
            B.last_its = B.h_ctl->it;
            c->adj_iterations += B.h_ctl->it;
            break;
        }
        if (B.h_ctl->done == 2)
            return afail(c, CMX_E_DIVERGED, "adjoint CG did not converge in " + std::to_string(kAdjMaxIts) +
                                                 " iterations");
        if (B.h_ctl->done == 3) return afail(c, CMX_E_DIVERGED, "adjoint CG: non-finite residual");
        chunk = 4;
    }
    adj_spmv_kernel<Rows><<<c->G, kCgBlock, 0, c->stream>>>(R, c->m64, c->n, B.cg.x, B.lz);
    ADJ_HIP(c, hipGetLastError());
    return CMX_OK;
}

// Slot layout of a stored step: V at [0, ps), the S-1 state planes after it.
int save_slot(cmx_ctx* c, double* dst)
{
    ADJ_HIP(c, hipMemcpyAsync(dst, c->V, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(dst + c->ps, c->state, (size_t)(c->S - 1) * c->ps * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    return CMX_OK;
}

int restore_slot(cmx_ctx* c, const double* src)
{
    ADJ_HIP(c, hipMemcpyAsync(c->V, src, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(c->state, src + c->ps, (size_t)(c->S - 1) * c->ps * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    c->stepped = false;
    return CMX_OK;
}

// Checkpoint plan: n_ck checkpoints every L steps, then per segment either
// its L traced steps (q = 0) or checkpoints every q steps and q traced steps
// at a time; q minimises the stored states when one level does not fit.
struct AdjPlan {
    int n_ck, q, slots;
};

AdjPlan adjoint_plan(int n_steps, int L, double slot_bytes, double limit)
{
    AdjPlan p{(n_steps + L - 1) / L, 0, 0};
    p.slots = p.n_ck + L + 1;
    if (p.slots * slot_bytes <= limit || L <= 2) return p;
    int best = p.slots, best_q = 0;
    for (int q = 1; q < L; ++q) {
        const int slots = p.n_ck + (L + q - 1) / q + q + 1;
        if (slots < best) {
            best = slots;
            best_q = q;
        }
    }
    p.q = best_q;
    p.slots = best;
    return p;
}

template <class Rows>
int adjoint_impl(cmx_ctx* c, const Rows& R, double t0, int n_steps, int L, int n_obs, const int32_t* obs_nodes,
                 int every, const double* obs_values, double* loss, double* grad_params, double* grad_state0)
{
    const int n = c->n, ps = c->ps, S = c->S, P = c->P;
    const size_t slot = (size_t)S * ps;
    const int n_obs_steps = n_steps / every;
    double limit = c->adj_mem_limit;
    if (!(limit > 0)) {
        size_t free_b = 0, total_b = 0;
        ADJ_HIP(c, hipMemGetInfo(&free_b, &total_b));
        limit = (double)free_b - (double)(1ull << 30);
    }
    const AdjPlan plan = adjoint_plan(n_steps, L, (double)slot * sizeof(double), limit);
    if (plan.slots * (double)slot * sizeof(double) > limit)
        return afail(c, CMX_E_NOMEM, "adjoint checkpoints: " + std::to_string(plan.slots) + " stored states of " +
                                         std::to_string((double)slot * sizeof(double) / 1e9) +
                                         " GB exceed the memory bound of " + std::to_string(limit / 1e9) + " GB");
    // CMX_PROGRESS set: one line on stderr per forward pass and replayed piece
    const bool verbose = std::getenv("CMX_PROGRESS") != nullptr;
    const int n_ck = plan.n_ck;
    const int sq = plan.q;                   // 0: one level
    const int traced = sq > 0 ? sq : L;        // steps kept at a time
    c->adj_slots = plan.slots;
    c->adj_sub = sq;
    Scratch sc;
    int rc;
    double *ck, *sub = nullptr, *seg, *final_slot, *obs_v, *obs_d;
    int32_t *d_obs_nodes, *obs_slot;
    AdjBuffers B{};
    if ((rc = sc.alloc(c, &ck, (size_t)n_ck * slot))) return rc;
    if (sq > 0 && (rc = sc.alloc(c, &sub, (size_t)((L + sq - 1) / sq) * slot))) return rc;
    if ((rc = sc.alloc(c, &seg, (size_t)traced * slot))) return rc;
    if ((rc = sc.alloc(c, &final_slot, slot))) return rc;
    if ((rc = sc.alloc(c, &obs_v, (size_t)std::max(n_obs_steps, 1) * n_obs))) return rc;
    if ((rc = sc.alloc(c, &obs_d, (size_t)std::max(n_obs_steps, 1) * n_obs))) return rc;
    if ((rc = sc.alloc(c, &d_obs_nodes, n_obs))) return rc;
    if ((rc = sc.alloc(c, &obs_slot, n))) return rc;
    if ((rc = sc.alloc(c, &B.lam_v, n))) return rc;
    if ((rc = sc.alloc(c, &B.lam_y, (size_t)(S - 1) * ps))) return rc;
    if ((rc = sc.alloc(c, &B.lz, n))) return rc;
    if ((rc = sc.alloc(c, &B.grad, (size_t)std::max(P, 1) * ps))) return rc;
    PlainCg& cg = B.cg;
    if ((rc = sc.alloc(c, &cg.r, n))) return rc;
    if ((rc = sc.alloc(c, &cg.w, n))) return rc;
    if ((rc = sc.alloc(c, &cg.z, n))) return rc;
    if ((rc = sc.alloc(c, &cg.s, n))) return rc;
    if ((rc = sc.alloc_gathered(c, &cg.x, n))) return rc;   // zero: the first solve starts from 0
    if ((rc = sc.alloc_gathered(c, &cg.p, n))) return rc;
    if ((rc = sc.alloc_gathered(c, &cg.m[0], n))) return rc;
    if ((rc = sc.alloc_gathered(c, &cg.m[1], n))) return rc;
    for (int q = 0; q < 3; ++q)
        if ((rc = sc.alloc(c, &cg.part[q], (size_t)kMaxBlocks * 3))) return rc;
    if ((rc = sc.alloc(c, &cg.hist, kAdjMaxIts + kAdjFirstMax + 1))) return rc;
    if ((rc = sc.alloc(c, &cg.ctl, 2))) return rc;
    cg.rtol = c->rtol > 0.0 ? c->rtol : 1e-10;
    // each solve starts from the previous step's mu (consecutive right-hand
    // sides lambda_V^{n+1} vary smoothly between observation steps)
    cg.warm = 1;
    cg.max_its = kAdjMaxIts;
    B.b = B.lam_v;
    std::vector<int32_t> h_slot(n, -1);
    for (int o = 0; o < n_obs; ++o) h_slot[obs_nodes[o]] = o;
    if (n_obs > 0) ADJ_HIP(c, hipMemcpy(d_obs_nodes, obs_nodes, n_obs * sizeof(int32_t), hipMemcpyHostToDevice));
    ADJ_HIP(c, hipMemcpy(obs_slot, h_slot.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    if (n_obs_steps > 0 && n_obs > 0)
        ADJ_HIP(c, hipMemcpy(obs_d, obs_values, (size_t)n_obs_steps * n_obs * sizeof(double), hipMemcpyHostToDevice));
    ADJ_HIP(c, hipMemset(B.grad, 0, (size_t)std::max(P, 1) * ps * sizeof(double)));
    SolveCtl* h_ctl = nullptr;
    ADJ_HIP(c, hipHostMalloc((void**)&h_ctl, sizeof(SolveCtl), hipHostMallocDefault));
    struct PinnedFree {
        SolveCtl* p;
        ~PinnedFree() { (void)hipHostFree(p); }
    } pinned{h_ctl};
    B.h_ctl = h_ctl;

    const int go = (n_obs + kAdjBlock - 1) / kAdjBlock;

    // Forward: checkpoints every L steps and the observed voltages, stepping
    // from one of those points to the next in one call.
    for (int m = 0; m < n_steps;) {
        if (m % L == 0 && (rc = save_slot(c, ck + (size_t)(m / L) * slot))) return rc;
        const int next = std::min({(m / L + 1) * L, (m / every + 1) * every, n_steps});
        if ((rc = step_range(c, t0, n_steps, m, next))) return rc;
        m = next;
        if (n_obs > 0 && m % every == 0) {
            gather_kernel<<<go, kAdjBlock, 0, c->stream>>>(n_obs, d_obs_nodes, c->V,
                                                           obs_v + (size_t)(m / every - 1) * n_obs);
            ADJ_HIP(c, hipGetLastError());
        }
    }
    if ((rc = save_slot(c, final_slot))) return rc;
    if (verbose) {
        std::fprintf(stderr, "cmx_adjoint: forward pass of %d steps done; %d stored states (segment checkpoints every "
                             "%d steps)\n", n_steps, plan.slots, sq);
        std::fflush(stderr);
    }
    std::vector<double> h_obs((size_t)n_obs_steps * n_obs);
    if (!h_obs.empty())
        ADJ_HIP(c, hipMemcpyAsync(h_obs.data(), obs_v, h_obs.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    double L2 = 0.0;
    for (size_t k = 0; k < h_obs.size(); ++k) {
        const double e = h_obs[k] - obs_values[k];
        L2 += e * e;
    }
    *loss = L2;

    // Backward, one checkpoint segment at a time.
    const double* d_at_end = (n_steps % every == 0 && n_obs_steps > 0)
                                 ? obs_d + (size_t)(n_steps / every - 1) * n_obs : nullptr;
    const int gc = (n + kAdjBlock - 1) / kAdjBlock;
    adjoint_final_kernel<<<gc, kAdjBlock, 0, c->stream>>>(n, ps, S - 1, c->V, obs_slot, d_at_end, B.lam_v, B.lam_y);
    ADJ_HIP(c, hipGetLastError());
    CellAdjArgs a{};
    a.n = n;
    a.ps = ps;
    a.n_sub = c->n_sub;
    a.dt_ode = c->dt_ode;
    a.am = c->am;
    a.cm = c->cm;
    a.c = c->am * c->cm / c->dt;
    a.clock = run_clock(t0, c->dt, n_steps);
    a.params = (const double*)c->params;
    a.pdef = (const double*)c->pdef;
    a.n_param_planes = c->n_param_planes;
    a.stim_id = c->stim_id;
    a.stim = c->stim;
    a.lz = B.lz;
    a.lam_y = B.lam_y;
    a.lam_v = B.lam_v;
    a.grad = P > 0 ? B.grad : nullptr;
    a.obs_slot = obs_slot;
    // Replays steps [r0, r1) from the state the context holds, its cell
    // kernels writing each step's starting state into seg (plain launches: a
    // captured step has no trace), then walks them backwards.
    auto replay_and_walk = [&](int r0, int r1) -> int {
        c->trace = seg;
        c->trace_k0 = r0;
        c->trace_len = r1 - r0;
        c->trace_slot = slot;
        const int graphs = c->use_graph;
        c->use_graph = 0;
        int rc2 = step_range(c, t0, n_steps, r0, r1);
        c->use_graph = graphs;
        c->trace = nullptr;
        if (rc2) return rc2;
        for (int m = r1 - 1; m >= r0; --m) {
            if ((rc2 = adjoint_solve<Rows>(c, R, B))) return rc2;
            const double* st = seg + (size_t)(m - r0) * slot;
            a.Vn = st;
            a.Yn = st + ps;
            a.k = m;
            a.obs_d = (m > 0 && m % every == 0) ? obs_d + (size_t)(m / every - 1) * n_obs : nullptr;
            model_ops(c->model)->adjoint(c, a, c->integrator == CMX_RUSH_LARSEN);
            ADJ_HIP(c, hipGetLastError());
        }
        if (verbose) {
            ADJ_HIP(c, hipStreamSynchronize(c->stream));
            std::fprintf(stderr, "cmx_adjoint: steps %d..%d back-propagated\n", r0, r1 - 1);
            std::fflush(stderr);
        }
        return CMX_OK;
    };
    for (int s0 = (n_ck - 1) * L; s0 >= 0; s0 -= L) {
        const int s1 = std::min(s0 + L, n_steps);
        if ((rc = restore_slot(c, ck + (size_t)(s0 / L) * slot))) return rc;
        if (sq == 0) {
            if ((rc = replay_and_walk(s0, s1))) return rc;
            continue;
        }
        // two levels: the segment once more, saving its state every q steps,
        // then each q-step piece replayed from its checkpoint, last first
        for (int m = s0; m < s1; m += sq) {
            if ((rc = save_slot(c, sub + (size_t)((m - s0) / sq) * slot))) return rc;
            if ((rc = step_range(c, t0, n_steps, m, std::min(m + sq, s1)))) return rc;
        }
        for (int r0 = s0 + (s1 - 1 - s0) / sq * sq; r0 >= s0; r0 -= sq) {
            if ((rc = restore_slot(c, sub + (size_t)((r0 - s0) / sq) * slot))) return rc;
            if ((rc = replay_and_walk(r0, std::min(r0 + sq, s1)))) return rc;
        }
    }
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    if (grad_params && P > 0) {
        std::vector<double> tmp((size_t)P * ps);
        ADJ_HIP(c, hipMemcpy(tmp.data(), B.grad, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int q = 0; q < P; ++q) std::copy(tmp.begin() + (size_t)q * ps, tmp.begin() + (size_t)q * ps + n,
                                              grad_params + (size_t)q * n);
    }
    if (grad_state0) {
        ADJ_HIP(c, hipMemcpy(grad_state0, B.lam_v, n * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<double> tmp((size_t)(S - 1) * ps);
        ADJ_HIP(c, hipMemcpy(tmp.data(), B.lam_y, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int s = 1; s < S; ++s) std::copy(tmp.begin() + (size_t)(s - 1) * ps, tmp.begin() + (size_t)(s - 1) * ps + n,
                                              grad_state0 + (size_t)s * n);
    }
    // leave the context at the end of the run, as cmx_step(ctx, t0, n_steps) would
    return restore_slot(c, final_slot);
}

}  // namespace
}  // namespace cmx

using namespace cmx;

extern "C" {

int cmx_adjoint(cmx_ctx* c, double t0, int32_t n_steps, int32_t checkpoint_every, int32_t n_obs,
                const int32_t* obs_nodes, int32_t obs_every, const double* obs_values, double* loss,
                double* grad_params, double* grad_state0)
{
    if (!c) return CMX_E_ARG;
    if (!loss) return afail(c, CMX_E_ARG, "null loss pointer");
    if (n_steps <= 0 || checkpoint_every <= 0 || obs_every <= 0)
        return afail(c, CMX_E_ARG, "n_steps, checkpoint_every and obs_every must be positive");
    if (n_obs < 0 || (n_obs > 0 && (!obs_nodes || !obs_values)))
        return afail(c, CMX_E_ARG, "observation nodes and values are required");
    for (int o = 0; o < n_obs; ++o)
        if (obs_nodes[o] < 0 || obs_nodes[o] >= c->n) return afail(c, CMX_E_ARG, "observation node out of range");
    for (int o = 0; o < n_obs; ++o)
        for (int o2 = 0; o2 < o; ++o2)
            if (obs_nodes[o2] == obs_nodes[o]) return afail(c, CMX_E_ARG, "observation nodes must be distinct");
    if (c->precision != CMX_FP64) return afail(c, CMX_E_ARG, "the adjoint needs precision CMX_FP64");
    if (!c->cell_groups.empty()) return afail(c, CMX_E_ARG, "the adjoint needs a single cell type");
    if (!model_ops(c->model) || !model_ops(c->model)->adjoint)
        return afail(c, CMX_E_ARG, "this cell model has no adjoint (O'Hara-Rudy, Maleckar 2008)");
    if (c->integrator != CMX_FORWARD_EULER && c->integrator != CMX_RUSH_LARSEN)
        return afail(c, CMX_E_ARG, "the adjoint supports forward-Euler and Rush-Larsen cells");
    if (c->split || c->svi) return afail(c, CMX_E_ARG, "the adjoint supports the ICI step only (no splitting, no SVI)");
    if (c->extrapolate) return afail(c, CMX_E_ARG, "the adjoint needs the reference's initial guess (extrapolate = 0)");
    if (c->act_every > 0) return afail(c, CMX_E_STATE, "activation tracking must be off during cmx_adjoint");
    if (c->format == CMX_FMT_STENCIL) {
        if (c->width == 4) return adjoint_impl(c, adj_stencil_rows<4>(c), t0, n_steps, checkpoint_every, n_obs,
                                               obs_nodes, obs_every, obs_values, loss, grad_params, grad_state0);
        if (c->width == 8) return adjoint_impl(c, adj_stencil_rows<8>(c), t0, n_steps, checkpoint_every, n_obs,
                                               obs_nodes, obs_every, obs_values, loss, grad_params, grad_state0);
        return adjoint_impl(c, adj_stencil_rows<16>(c), t0, n_steps, checkpoint_every, n_obs, obs_nodes, obs_every,
                            obs_values, loss, grad_params, grad_state0);
    }
    const EllRows R = ell_rows(c);
    return adjoint_impl(c, R, t0, n_steps, checkpoint_every, n_obs, obs_nodes, obs_every, obs_values, loss,
                        grad_params, grad_state0);
}

int cmx_set_adjoint_memory(cmx_ctx* c, double bytes)
{
    if (!c) return CMX_E_ARG;
    if (!(bytes >= 0)) return afail(c, CMX_E_ARG, "the adjoint memory bound must be >= 0");
    c->adj_mem_limit = bytes;
    return CMX_OK;
}

}  // extern "C"
