// Batched single-cell kernels (cmx_cells_run): one cell per thread, or two
// wavefronts per 64 cells.  Templates on the model, instantiated by each
// model's translation unit (cmx_model_*.hip).
#pragma once
#include <type_traits>

#include "cmx_context.h"

namespace cmx {

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
    // gate lookup table (GateCols layout), rows at lut_vmin + r / lut_inv_dv, or null
    const float* __restrict__ lut;
    float lut_vmin, lut_inv_dv;
    int lut_rows;
    int32_t* __restrict__ lut_miss;        // 1 + a cell whose voltage fell outside the table
    // V samples, or null: trace[s * trace_n + i] = V after step s * trace_every
    // (s = 0: the initial V)
    double* __restrict__ trace;
    int trace_every, trace_n;
};

template <class Model, class RS, int IG, bool LUT = false>
__global__ __launch_bounds__(kBlock) void cells_kernel(CellsArgs<RS> a, int k0, int k1)
{
    constexpr int NS = Model::NS;
    constexpr int NPd = Model::NP > 0 ? Model::NP : 1;
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    static_assert(!LUT || RL, "lookup tables serve Rush-Larsen cells");
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.trace && k0 == 0) a.trace[i] = a.V[i];
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
        // with a table the gate outputs above are unused (and not computed)
        float tgi[NS], tgd[NS];
        if constexpr (LUT) {
            if (!lut_gates<Model>(a.lut, a.lut_vmin, a.lut_inv_dv, a.lut_rows, (float)V, tgi, tgd)) *a.lut_miss = i + 1;
        }
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            if (Model::gate(s)) {
                if (LUT) y[s] = tgi[s] + (y[s] - tgi[s]) * tgd[s];
                else if (RL) y[s] = gi[s] + (y[s] - gi[s]) * cmx_decay(gk[s], h);
                else y[s] += h * dy[s];
            } else {
                yd[s] += a.dt * (double)dy[s];
                y[s] = (RS)yd[s];
            }
        }
        const double v0 = vd;
        vd += a.dt * (double)dv;
        if (a.trace && (k + 1) % a.trace_every == 0) a.trace[(size_t)((k + 1) / a.trace_every) * a.trace_n + i] = vd;
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

// Two wavefronts per 64 cells (lanes_per_cell = 2): wavefront 0 of the block
// advances the cells' gates, wavefront 1 everything else (currents, the
// non-gate states, V and the CellProperties summary).  A gate's Rush-Larsen
// (or Euler) coefficients depend on V only, so the first wavefront needs only
// V from the second, and the second needs the gates from the first: they
// exchange both through LDS once per step (one barrier; buffers alternate
// with the step's parity).  Each wavefront runs the model's right-hand side
// and uses only its own outputs, so the other part is dead code it does not
// execute.  The step's arithmetic is the one-lane kernel's, split: a cell's
// critical path per step is the longer of the two parts, which is what bounds
// a sweep whose wavefronts are fewer than the SIMDs can interleave.
constexpr int kSplitCells = 64;

template <class Model, class RS, int IG>
__global__ __launch_bounds__(2 * kSplitCells) void cells_split_kernel(CellsArgs<RS> a, int k0, int k1)
{
    constexpr int NS = Model::NS;
    constexpr int NG = GateCols<Model>::count();
    constexpr int NPd = Model::NP > 0 ? Model::NP : 1;
    constexpr bool RL = IG == CMX_RUSH_LARSEN;
    __shared__ RS s_gate[2][NG > 0 ? NG : 1][kSplitCells];
    __shared__ double s_v[2][kSplitCells];
    const int lane = threadIdx.x & (kSplitCells - 1);
    const bool gates_role = threadIdx.x < kSplitCells;
    const int j = blockIdx.x * kSplitCells + lane;
    const bool valid = j < a.n;
    const int i = valid ? j : a.n - 1;
    // the block's longest cell bounds its steps (both wavefronts hold the same
    // 64 cells, so each reduces the same maximum): every thread of the block
    // runs the same number of barriers
    const int kend = valid ? min(k1, a.n_steps[i]) : k0;
    int kmax = kend;
#pragma unroll
    for (int off = kSplitCells / 2; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off, kSplitCells));
    if (k0 >= kmax) return;
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
    if (a.trace && k0 == 0 && valid && !gates_role) a.trace[i] = vd;
    for (int k = k0; k < kmax; ++k) {
        const int b = k & 1;
        if (k < kend) {
            const double t = k * a.dt;
            const RS V = (RS)vd;
            y[0] = V;
            RS dy[NS], gi[NS], gk[NS];
            if (gates_role) {
                Model::template eval<RS, RL>(y, V, RS(0), p, cm, dy, gi, gk);
#pragma unroll
                for (int s = 1; s < NS; ++s) {
                    if (!Model::gate(s)) continue;
                    if (RL) y[s] = gi[s] + (y[s] - gi[s]) * cmx_decay(gk[s], h);
                    else y[s] += h * dy[s];
                    s_gate[b][GateCols<Model>::col(s)][lane] = y[s];
                }
            } else {
                const RS is = (RS)(a.stim ? stimulus_at(st, t) : 0.0);
                Model::template eval<RS, RL>(y, V, is, p, cm, dy, gi, gk);
                const RS dv = Model::template dvdt<RS>(y, V, is, p, cm);
#pragma unroll
                for (int s = 1; s < NS; ++s) {
                    if (Model::gate(s)) continue;
                    yd[s] += a.dt * (double)dy[s];
                    y[s] = (RS)yd[s];
                }
                const double v0 = vd;
                vd += a.dt * (double)dv;
                s_v[b][lane] = vd;
                if (a.trace && (k + 1) % a.trace_every == 0)
                    a.trace[(size_t)((k + 1) / a.trace_every) * a.trace_n + i] = vd;
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
        }
        __syncthreads();
        if (k < kend) {
            if (gates_role) {
                vd = s_v[b][lane];
            } else {
#pragma unroll
                for (int s = 1; s < NS; ++s)
                    if (Model::gate(s)) y[s] = s_gate[b][GateCols<Model>::col(s)][lane];
            }
        }
    }
    if (!valid) return;
    if (gates_role) {
#pragma unroll
        for (int s = 1; s < NS; ++s)
            if (Model::gate(s)) a.state[(size_t)(s - 1) * a.ps + i] = (double)y[s];
        return;
    }
#pragma unroll
    for (int s = 1; s < NS; ++s)
        if (!Model::gate(s)) a.state[(size_t)(s - 1) * a.ps + i] = yd[s];
    a.V[i] = vd;
    a.phase[i] = phase;
    a.ap_count[i] = count;
    a.onset[i] = onset;
    a.peak[i] = peak;
    a.last_peak[i] = last_peak;
    a.end[i] = end;
}

template <class Model, class RS, int IG>
void launch_cells(const CellsArgs<RS>& a, int k0, int k1, hipStream_t s, int lanes = 1)
{
    if (lanes == 2) {
        cells_split_kernel<Model, RS, IG><<<(a.n + kSplitCells - 1) / kSplitCells, 2 * kSplitCells, 0, s>>>(a, k0, k1);
        return;
    }
    if constexpr (IG == CMX_RUSH_LARSEN && std::is_same<RS, float>::value) {
        if (a.lut) {
            cells_kernel<Model, RS, IG, true><<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a, k0, k1);
            return;
        }
    }
    cells_kernel<Model, RS, IG><<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a, k0, k1);
}

}  // namespace cmx
