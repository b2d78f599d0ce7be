#!/usr/bin/env python3
"""Benchmark: monodomain steps/s on one MI355X, FP32-mixed vs FP64.

Workload (the reference's 3-D slab case): ConstructRegularSlabMesh(0.01, 2.0,
0.7, 0.3) -> 201 x 71 x 31 = 442,401 nodes, 2,520,000 tetrahedra; ten
Tusscher 2006 epicardial cells with Rush-Larsen; sigma = diag(1.75, 0.19, 0.19)
mS/cm with fibres along x; Am 1400 /cm, Cm 1 uF/cm^2; dt_pde = dt_ode =
0.01 ms; stimulus -50000 uA/cm^3 for 1 ms on x < 0.03 cm; Jacobi-PCG rtol
1e-8.  A step is one full PDE step (cell sweep + ICI right-hand side + CG
solve) with every input resident in HBM.

value = FP32-mixed steps/s (FP32 cells/CG with FP64 voltage, dots and
residual refinement); the FP64 path is timed in the same run, and the
accuracy of the mixed run against it is reported (max / RMS dV after the
same steps).  The CPU baseline is the FP64 oracle (the same algorithm in
C++/OpenMP) on this host's cores over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

FORMAT_NAMES = {1: 'stencil', 2: 'ell', 3: 'ell_offsets'}

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EX, EY, EZ, H = 200, 70, 30, 0.01
SIGMA = (1.75, 0.19, 0.19)
STIM = (-50000.0, 0.0, 1.0)       # magnitude uA/cm^3, start ms, duration ms
STIM_PROPAGATING = -100000.0      # a magnitude above threshold for the second case
STIM_XLIM = 0.03
DT = 0.01
EXTRAPOLATE = 2      # CG initial guess V^n + 2 x_{n-1} - x_{n-2} (corrections x), in both paths
ACC_STEPS = 5000     # accuracy runs: the whole 50 ms
ORACLE_STEPS = 200   # ... with the voltage after this many steps checked against the oracle


USE_GRAPH = True
AUTOTUNE = True      # cmx_autotune picks the solver block count after the warmup
LOOKUP_TABLES = False  # mixed path: gate coefficients from a table in V


def build_problem(precision, coords, elems, stim_index, cmx, stim_mag=STIM[0], tables=False):
    return cmx.Monodomain(coords, elems, np.array(SIGMA), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                          stimuli=[(stim_mag, STIM[1], STIM[2])], stim_index=stim_index, precision=precision,
                          dt_pde=DT, dt_ode=DT, ksp_rtol=1e-8, ksp_atol=0.0, ksp_max_its=500,
                          refine_sweeps=1 if precision == cmx.MIXED else 0, use_graph=USE_GRAPH,
                          extrapolate=EXTRAPOLATE, lookup_tables=tables)


def bytes_per_row(kind, precision, n_states, refined, guess=EXTRAPOLATE):
    """Algorithmic HBM bytes per row of one launch of a kernel kind that does
    its work.  s = CG-vector / forcing size, ss = cell-state size; V and Z are
    FP64; the pattern id and the stimulus id are one byte each.
      cell      non-V states in+out, stimulus id, F out; settling the previous
                step: Z, x1 (+ x2) in, V out; with an initial guess x0 in/out,
                Y = V + x0 and G = F - c x0 out (+ x_{n-2} in/out at order 2)
      residual  Z, G gathered (read once), pattern id in; r, u0, x = 0 out,
                + FP64 r0 when refined; the refinement sweep reads r0, x1 and
                writes r, u, x2 (averaged over the two launches of a refined step)
      spmv      (solve start) u0 gathered, r in; w, m out
      update    (pipelined CG iteration) m gathered, r w x z s p in; z s p x r w m out
      finish    (batch end) Z, x1 (+ x2) in, V out"""
    s = 4 if precision == 1 else 8
    ss = s
    if kind == 'cell':
        b = 2 * (n_states - 1) * ss + 1 + s                  # states, stimulus id, F
        b += 8 + s + (s if refined else 0) + 8               # Z, x1, x2 in; V out
        if guess:
            b += 2 * s + 8 + s                               # x0 in/out, Y, G out
            if guess >= 2:
                b += 2 * s                                   # x_{n-2} in/out
        return b
    if kind == 'residual':
        first = 8 + s + 1 + 3 * s + (8 if refined else 0)
        if not refined:
            return first
        return (first + (8 + s + 1 + 3 * s)) / 2
    if kind == 'spmv':
        return 4 * s + 1
    if kind == 'update':
        return 14 * s + 1
    if kind == 'finish':
        return 16 + s * (2 if refined else 1)
    return 0


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


KINDS = ['cell', 'residual', 'spmv', 'update', 'finish']


def progress(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def time_path(m, steps, warmup, roof_steps, events=True):
    """W untimed steps from t = 0, then exactly K timed steps with no
    instrumentation (value); then an untimed 5-step pass with HIP events on
    every kernel kind ranks them, and roof_steps more steps with events
    around every launch of the dominant kind only (its average launch
    duration for the roofline)."""
    t = 0.0
    if warmup:
        m.step(t, warmup)
        t += warmup * DT
    tuned = None
    if AUTOTUNE:
        # solver block count picked from the warmed-up state (state restored)
        chosen, times = m.autotune(t, 64)      # one batch of steps per trial
        tuned = {'blocks': chosen, 'ms_per_step': {str(k): round(v, 5) for k, v in times.items()},
                 'chosen': m.launch_config()}
        progress('autotune: %s (blocks %s)' % (tuned['chosen'], tuned['ms_per_step']))
    st0 = m.stats()
    sync()
    t0 = time.perf_counter()
    m.step(t, steps)
    sync()
    elapsed = time.perf_counter() - t0
    progress('timed %d steps: %.3f s' % (steps, elapsed))
    t += steps * DT
    st1 = m.stats()
    its = (st1['cg_iterations'] - st0['cg_iterations']) / steps
    ref = (st1['refine_sweeps'] - st0['refine_sweeps']) / steps
    graphs = {'mode': ['plain launches', 'graph replay', 'plain launches (profiler attached)'][st1['graph_mode']],
              'captures_in_timed_region': st1['graph_captures'] - st0['graph_captures'],
              'replays_in_timed_region': st1['graph_launches'] - st0['graph_launches'],
              'resumes_in_timed_region': st1['resumes'] - st0['resumes'],
              'folded_refinements_in_timed_region': st1['folded_refinements'] - st0['folded_refinements']}
    if not events:
        return {'elapsed': elapsed, 'its': its, 'refine': ref, 'ranking': {}, 'rank_its': its,
                'dominant': 'cell', 'dominant_ms': float('nan'), 'dominant_launches': 1, 'roof_its': its,
                'autotune': tuned, 'graphs': graphs}
    m.set_profiling(KINDS)
    m.kernel_times(reset=True)
    m.step(t, 5)
    t += 5 * DT
    ranking = {k: v for k, v in m.kernel_times(reset=True).items() if k in KINDS}
    dominant = max(ranking, key=lambda k: ranking[k][0])
    st_r = m.stats()
    m.set_profiling([dominant])
    m.kernel_times(reset=True)
    m.step(t, roof_steps)
    st2 = m.stats()
    ktimes = m.kernel_times(reset=True)
    m.set_profiling([])
    rank_its = (st_r['cg_iterations'] - st1['cg_iterations']) / 5
    roof_its = (st2['cg_iterations'] - st_r['cg_iterations']) / roof_steps
    return {'elapsed': elapsed, 'its': its, 'refine': ref, 'ranking': ranking, 'rank_its': rank_its,
            'dominant': dominant, 'dominant_ms': ktimes[dominant][0], 'dominant_launches': ktimes[dominant][1],
            'roof_its': roof_its, 'autotune': tuned, 'graphs': graphs}


def working_launches_per_step(kind, its, refined, launches_per_step):
    """Launches per step that do their work: every iteration launch past a
    solve's convergence exits at once; the finish kernel runs once per
    cmx_step batch, not once per step."""
    if kind == 'update':
        return its
    if kind == 'spmv':
        return 1     # the first solve's start; a refinement solve rarely iterates
    if kind == 'residual':
        return 2 if refined else 1
    if kind == 'finish':
        return launches_per_step
    return 1


def survey_bytes(n, nnz, n_states, precision, its, refine):
    """The survey's algorithmic bytes per PDE step (Jacobi-PCG textbook
    minimum; s = 8 in FP64, 4 in the mixed path): cells n (8 V + 8 Iion +
    2 S s + 4 stimulus), right-hand side nnz (s + 4) + 4 (n + 1) + 24 n, one
    CG iteration nnz (s + 4) + 4 (n + 1) + 18 n s, one FP64 refinement sweep
    12 nnz + 4 (n + 1) + 48 n."""
    s = 4 if precision == 1 else 8
    S = n_states - 1
    cell = n * (16 + 2 * S * s + 4)
    rhs = nnz * (s + 4) + 4 * (n + 1) + 24 * n
    it = nnz * (s + 4) + 4 * (n + 1) + 2 * n * s + 13 * n * s + 3 * n * s
    ref = 12 * nnz + 4 * (n + 1) + 48 * n
    return {'cell': cell, 'rhs': rhs, 'cg_iteration': it, 'refinement': ref,
            'step': cell + rhs + its * it + refine * ref}


def cpu_baseline(n_steps, threads):
    from oracle import pyoracle
    pyoracle.lib()
    r = pyoracle.mono_run(3, EX, EY, EZ, H, model=pyoracle.TT06_EPI, rush_larsen=True, sigma=SIGMA,
                          stim_mag=STIM[0], stim_dur=STIM[2], stim_start=STIM[1], stim_xlim=STIM_XLIM,
                          t_end=n_steps * DT, ksp_rtol=1e-8, threads=threads)
    s = r['stats']
    loop = s['t_ode'] + s['t_rhs'] + s['t_ksp']
    return s['steps'] / loop, s


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def oracle_slab_run(stim_mag, n_steps, threads):
    """The oracle's FP64 slab run (V^n warm start, Jacobi-PCG rtol 1e-8) to
    step n_steps: the accuracy leg's reference for the first steps."""
    from oracle import pyoracle
    return pyoracle.mono_run(3, EX, EY, EZ, H, model=pyoracle.TT06_EPI, rush_larsen=True, sigma=SIGMA,
                             stim_mag=stim_mag, stim_dur=STIM[2], stim_start=STIM[1], stim_xlim=STIM_XLIM,
                             t_end=n_steps * DT, ksp_rtol=1e-8, threads=threads)['V']


def cfg1_case(cmx, steps=1000):
    """The reference's CPU-runnable case: LR91 FE on the 101 x 101 staggered
    square (h 0.01 cm), sigma 1.75, -25000 uA/cm^3 for 0.5 ms on x < 0.02,
    10 ms = 1000 steps, CG rtol 1e-8; launch-latency bound at 10,201 nodes.
    Both paths, 1000 steps after a 20-step warmup on the same context, then
    the mixed voltage against the FP64 one after the same 1000 steps from
    t = 0."""
    coords, elems = cmx.slab_mesh(2, 100, 100, 0, H)
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    init = np.repeat(np.asarray(cmx.initial_state(cmx.LR91), dtype=np.float64)[:, None], coords.shape[0], axis=1)
    out, V = {}, {}
    for name, prec in (('fp64', cmx.FP64), ('mixed', cmx.MIXED)):
        m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.FORWARD_EULER,
                           stimuli=[(-25000.0, 0.0, 0.5)], stim_index=stim_index, precision=prec, ksp_rtol=1e-8,
                           ksp_atol=0.0, refine_sweeps=1 if prec == cmx.MIXED else 0, use_graph=USE_GRAPH)
        m.step(0.0, 20)
        m.step(0.2, steps)     # graphs of the budgets this run uses
        m.set_state(init)
        m.step(0.0, 20)
        sync()
        st0 = m.stats()
        t = time.perf_counter()
        m.step(0.2, steps)
        sync()
        el = time.perf_counter() - t
        st1 = m.stats()
        m.set_state(init)
        m.step(0.0, steps)
        V[name] = m.voltage()
        out[name] = {'steps_per_s': round(steps / el, 1), 'ms_per_step': round(1e3 * el / steps, 4),
                     'cg_its_per_step': round((st1['cg_iterations'] - st0['cg_iterations']) / steps, 3),
                     'graph_captures_in_timed_region': st1['graph_captures'] - st0['graph_captures']}
        m.close()
    dv = np.abs(V['mixed'] - V['fp64'])
    return {'workload': 'LR91 FE monodomain square 101x101 nodes h=0.01, stim -25000 uA/cm^3 0.5 ms x<0.02, '
                        'dt 0.01 ms, Jacobi-CG rtol 1e-8, %d timed steps after 20' % steps,
            'n_nodes': int(coords.shape[0]), 'paths': out,
            'mixed_over_fp64': round(out['mixed']['steps_per_s'] / out['fp64']['steps_per_s'], 3),
            'max_dV_mixed_vs_fp64_mV': float(dv.max())}


def adjoint_case(args):
    """The adjoint workload: a 2-D LR91 square of 256 x 256 nodes (h 0.01 cm,
    forward Euler, FP64, Jacobi-PCG rtol 1e-12), per-node scalings of g_Na and
    g_Kmax; ground truth = seeded smooth Gaussian random fields in [0.6, 1.4]
    with correlation length 0.1 cm; loss = squared V mismatch at 64 seeded
    observation nodes every 0.1 ms over 20 ms (2000 steps) against the
    ground-truth run, evaluated at scalings 1; checkpoints every 100 steps.
    Prints one JSON line: the gradient's wall time next to the forward run's,
    and its agreement with FP64 central differences of the device forward run
    along random directions."""
    from chaste_amd import cmx
    from chaste_amd.autograd import smooth_random_field
    N, steps, every, n_obs = 255, 2000, 10, 64
    coords, elems = cmx.slab_mesh(2, N, N, 0, H)
    n = coords.shape[0]
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.FORWARD_EULER,
                       stimuli=[(-400000.0, 0.0, 0.5)], stim_index=stim_index, precision=cmx.FP64,
                       ksp_rtol=1e-12, ksp_atol=0.0, use_graph=USE_GRAPH)
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(cmx.LR91)])
    rng = np.random.default_rng(2024)
    truth = np.stack([smooth_random_field((N + 1, N + 1), H, 0.1, rng).ravel() for _ in range(2)])
    obs = np.sort(rng.choice(n, n_obs, replace=False)).astype(np.int32)

    def params(s):
        p = np.repeat(defaults[:, None], n, axis=1)
        p[:2] *= s
        return p

    def observe(s):
        m.set_state(y0)
        m.set_params(params(s))
        out = []
        for j in range(steps // every):
            m.step(j * every * DT, every)
            out.append(m.voltage()[obs])
        return np.array(out)

    progress('adjoint: ground truth')
    data = observe(truth)
    s1 = np.ones((2, n))
    m.set_state(y0)
    m.set_params(params(s1))
    m.step(0.0, 50)
    sync()
    m.set_state(y0)
    t = time.perf_counter()
    m.step(0.0, steps)
    sync()
    t_fwd = time.perf_counter() - t
    progress('adjoint: gradient')
    m.set_state(y0)
    t = time.perf_counter()
    it0 = m.stats()['adjoint_iterations']
    loss, gp, _ = m.adjoint(0.0, steps, obs, every, data, 100)
    t_adj = time.perf_counter() - t
    adj_its = m.stats()['adjoint_iterations'] - it0
    progress('adjoint: %.3f s, %d CG iterations (%.1f per step)' % (t_adj, adj_its, adj_its / steps))
    g = gp[:2] * defaults[:2, None]

    def central(d, eps):
        return (((observe(s1 + eps * d) - data) ** 2).sum() - ((observe(s1 - eps * d) - data) ** 2).sum()) / (2 * eps)

    # Central differences at eps and eps/2, Richardson-combined.  LR91's h and
    # j gate rates jump at V = -40 mV, so the loss jumps wherever a
    # perturbation moves some node's voltage at some step across -40 mV; the
    # two differences then disagree far beyond their O(eps^2) gap, and the
    # direction is differenced again at a 100 times smaller eps.
    rel, retried = [], 0
    for k in range(args.adjoint_directions):
        d = rng.standard_normal((2, n))
        ad = float((g * d).sum())
        for eps in (1e-4, 1e-6):
            d1, d2 = central(d, eps), central(d, eps / 2)
            if abs(d1 - d2) <= 1e-5 * abs(d2):
                break
            retried += eps == 1e-4
        fd = (4 * d2 - d1) / 3
        rel.append(abs(ad - fd) / abs(fd))
        progress('adjoint: direction %d eps %g rel %.2e' % (k, eps, rel[-1]))
    line = {
        'metric': 'adjoint gradient wall time (loss + dL/d 2x65536 conductance scalings)',
        'value': round(t_adj, 4), 'unit': 's', 'higher_is_better': False, 'n_gpus': 1, 'dtype': 'f64',
        'data': 'synthetic (seeded Gaussian random-field scalings, seeded observation nodes)',
        'config': {'workload': 'LR91 FE monodomain square 256x256 nodes h=0.01, 20 ms (2000 steps), '
                               'stim -400000 uA/cm^3 0.5 ms x<0.02, Jacobi-CG rtol 1e-12, 64 observation '
                               'nodes every 0.1 ms, checkpoints every 100 steps', 'n_nodes': n},
        'forward_s': round(t_fwd, 4), 'gradient_over_forward': round(t_adj / t_fwd, 2), 'loss': loss,
        'adjoint_cg_iterations_per_step': round(adj_its / steps, 2),
        'fd_directions': len(rel), 'fd_eps': [1e-4, 1e-6],
        'fd_directions_redifferenced': retried,
        'fd_within_1e-4': int(sum(r <= 1e-4 for r in rel)),
        'fd_max_rel_err': max(rel) if rel else None,
        'fd_median_rel_err': float(np.median(rel)) if rel else None,
        'fd_note': 'Richardson-combined central differences of the device forward run at eps and eps/2; a '
                   'direction whose two differences disagree (the loss jumps where a node crosses LR91\'s '
                   'h/j-gate branch at -40 mV) is differenced again at eps 1e-6',
    }
    print(json.dumps(line))


def cells_case(args):
    """The batched-cell workload: 65,536 independent O'Hara-Rudy 2011
    endocardial cells (41 states; Rush-Larsen, dt 0.005 ms, FP32 rates and
    gates with FP64 V and concentrations), per-cell scalings of the seven
    conductances G_Na, P_CaL, G_Kr, G_Ks, G_K1, G_to, G_NaL log-uniform in
    [0.5, 2], cycle lengths uniform in 300-1000 ms with a 2 ms -80 uA/uF
    stimulus at the start of each cycle, 2-10 cycles per cell (uneven run
    lengths), all seeded (--cells-model tt06 runs the same sweep on ten
    Tusscher 2006 cells with its six conductances).  Prints one JSON line:
    cell-steps/s over the launches, and the mixed run's accuracy against the
    FP64 device run on a seeded subset."""
    from chaste_amd import cmx
    model = cmx.ORD_ENDO if args.cells_model == 'ord' else cmx.TT06_EPI
    name = {cmx.ORD_ENDO: "O'Hara-Rudy 2011 endo (41 states)", cmx.TT06_EPI: 'TT06-epi (19 states)'}[model]
    n = args.cells
    rng = np.random.default_rng(355)
    defaults = np.array([d for _, d in cmx.param_info(model)])
    params = defaults[:, None] * np.exp(rng.uniform(np.log(0.5), np.log(2.0), (len(defaults), n)))
    period = np.round(rng.uniform(300.0, 1000.0, n))
    beats = rng.integers(2, 11, n)
    t_end = period * beats
    stimuli = [(-80.0, 0.0, 2.0, float(p)) for p in period]
    progress('cells: %d %s mixed' % (n, name))
    y, props, st = cmx.cells_run(model, cmx.RUSH_LARSEN, t_end, 0.005, stimuli=stimuli, params=params,
                                 precision=cmx.MIXED, threshold=-30.0, lookup_tables=LOOKUP_TABLES,
                                 lanes_per_cell=args.lanes_per_cell)
    sub = np.sort(rng.choice(n, min(args.cells_check, n), replace=False))
    progress('cells: %d FP64 for the accuracy check' % len(sub))
    y64, p64, st64 = cmx.cells_run(model, cmx.RUSH_LARSEN, t_end[sub], 0.005,
                                   stimuli=[stimuli[i] for i in sub], params=params[:, sub], precision=cmx.FP64,
                                   threshold=-30.0)
    steps = np.round(t_end / 0.005)
    sps = st['cell_steps'] / (st['kernel_ms'] * 1e-3)
    same = p64[0] == props[0][sub]
    line = {
        'metric': 'batched single cells: cell ODE steps/s (FP32 rates and gates, FP64 V and concentrations)',
        'value': round(sps, 1), 'unit': 'cell-steps/s', 'higher_is_better': True, 'n_gpus': 1,
        'dtype': 'f32 (f64 V, concentrations)', 'data': 'synthetic (seeded conductance scalings and pacing)',
        'config': {'workload': '%d %s cells, Rush-Larsen dt 0.005 ms, %d conductance scalings log-uniform '
                               '[0.5, 2], CL uniform 300-1000 ms, 2 ms -80 uA/uF stimulus per cycle, 2-10 cycles '
                               'per cell' % (n, name, len(defaults))
                               + (', gate lookup tables (0.02 mV rows)' if LOOKUP_TABLES else ''),
                   'n_cells': n, 'lookup_tables': LOOKUP_TABLES},
        'kernel_s': round(st['kernel_ms'] * 1e-3, 3), 'cell_steps': st['cell_steps'], 'launches': st['launches'],
        'lanes_per_cell': args.lanes_per_cell, 'split_launches': st['split_launches'],
        'longest_cell_steps': int(steps.max()),
        'lane_balance': round(float(steps.sum() / (n * steps.max())), 3),
        'all_lanes_busy_rate': round(n * float(steps.max()) / (st['kernel_ms'] * 1e-3), 1),
        'note': 'one cell per lane; the run ends with the longest cell (its steps are sequential), so the '
                'rate the lanes sustain while busy is value / lane_balance',
        'accuracy_mixed_vs_fp64': {
            'cells': len(sub), 'max_dV_mV': float(np.abs(y[0][sub] - y64[0]).max()),
            'ap_counts_equal': int(same.sum()),
            'onset_max_err_ms': float(np.abs(props[1][sub] - p64[1])[same].max()),
            'ap_end_max_err_ms': float(np.abs(props[3][sub] - p64[3])[same].max()),
            'fp64_kernel_s': round(st64['kernel_ms'] * 1e-3, 3)},
        'aps_completed': int(props[0].sum()),
    }
    print(json.dumps(line))


def shells_case(args):
    """The batched-tissue workload: 8 truncated prolate-ellipsoid shells
    (inner semi-axes 2.5/1.0 cm, outer 3.5/2.0 cm, cut at the base), meshed as
    tetrahedra at h = 0.04 ... 0.02 cm (0.52M - 4.0M nodes each, uneven;
    --shell-h sets the range),
    fibres rotating -60 -> +60 degrees through the wall, conductivities
    1.75/0.19 mS/cm, TT06-epi RL, -100000 uA/cm^3 for 1 ms on each apical rim.
    The shells form one context (a disjoint union: block-diagonal operators,
    one launch sequence per step).  FP64 and mixed run 20 ms with activation
    tracking; the mixed run then continues to 100 ms for its rate."""
    from chaste_amd import cmx
    from chaste_amd.shells import disjoint_union, ellipsoid_shell
    progress('shells: meshing')
    h_max, h_min = args.shell_h
    meshes = [ellipsoid_shell(h) for h in np.linspace(h_max, h_min, 8)]
    sizes = [int(m[0].shape[0]) for m in meshes]
    coords, elems, fibres, apex, _ = disjoint_union(meshes)
    del meshes
    sig = cmx.conductivity_tensors(fibres, (1.75, 0.19), axisymmetric=True)
    del fibres
    n = coords.shape[0]
    stim_index = np.where(apex, 0, -1).astype(np.int32)
    warm, first, total = 20, 2000, args.shell_steps
    out, runs = {}, {}
    for name, prec in (('fp64', cmx.FP64), ('mixed', cmx.MIXED)):
        progress('shells: %s setup (%d nodes, %d tetrahedra)' % (name, n, elems.shape[0]))
        t = time.perf_counter()
        m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                           stimuli=[(-100000.0, 0.0, 1.0)], stim_index=stim_index, precision=prec, ksp_rtol=1e-8,
                           ksp_max_its=500, refine_sweeps=1 if prec == cmx.MIXED else 0, sigma_elem=sig,
                           use_graph=USE_GRAPH, extrapolate=EXTRAPOLATE)
        setup = time.perf_counter() - t
        m.set_activation(0.0, 1, 0.0)
        m.step(0.0, warm)
        sync()
        st0 = m.stats()
        t = time.perf_counter()
        m.step(warm * DT, first - warm)
        sync()
        el = time.perf_counter() - t
        runs[name] = (m.voltage(), m.activation()[4])
        steps = first - warm
        common = {'steps_per_s': round(steps / el, 2), 'window_ms': [warm * DT, first * DT]}
        k = first
        while prec == cmx.MIXED and k < total:
            chunk = min(2000, total - k)
            t = time.perf_counter()
            m.step(k * DT, chunk)
            sync()
            el += time.perf_counter() - t
            steps += chunk
            k += chunk
            progress('shells: mixed at %g ms' % (k * DT))
        st1 = m.stats()
        # 20 more steps with HIP events on every launch (plain launches): each
        # kernel kind's time per step, and the ELL iteration's bandwidth over
        # the launches that iterated (the others exit on the converged test)
        m.set_profiling(KINDS)
        m.kernel_times(reset=True)
        m.step(k * DT, 20)
        kt = m.kernel_times(reset=True)
        m.set_profiling([])
        st2 = m.stats()
        s_b = 4 if prec == cmx.MIXED else 8
        w = st2['row_width']
        its = st2['cg_iterations'] - st1['cg_iterations']
        # vectors + one ELL row of A: a value per slot, and a 4-byte column per
        # slot (column-index ELL) or one set byte per row (offset-set ELL)
        row_b = 14 * s_b + w * s_b + (1 if st2['format'] == cmx.FMT_ELL_OFFSETS else 4 * w)
        upd_bytes = n * row_b
        kernels = {kk: {'ms_per_step': round(kt[kk][0] / 20, 4), 'launches_per_step': kt[kk][1] / 20}
                   for kk in KINDS if kt[kk][1]}
        if kt['update'][1] and its:
            kernels['update']['working_launch_GBps'] = round(upd_bytes * its / (kt['update'][0] * 1e-3) / 1e9, 1)
            kernels['update']['bytes_per_row'] = row_b
        out[name] = {'steps_per_s': round(steps / el, 2), 'node_steps_per_s': round(steps * n / el, 1),
                     'timed_steps': steps, 'timed_window_ms': [warm * DT, k * DT], 'setup_s': round(setup, 2),
                     'common_window': common,
                     'cg_its_per_step': round((st1['cg_iterations'] - st0['cg_iterations']) / steps, 3),
                     'format': FORMAT_NAMES[st1['format']], 'offset_sets': st1['n_patterns'],
                     'row_width': st1['row_width'], 'kernels_20_steps': kernels}
        m.close()
    n_elems = int(elems.shape[0])
    del coords, elems, sig, stim_index
    adj = shell_adjoint(h_max)
    dv = np.abs(runs['mixed'][0] - runs['fp64'][0])
    a64, a32 = runs['fp64'][1], runs['mixed'][1]
    both = (a64 >= 0) & (a32 >= 0)
    line = {
        'metric': 'batched tissue: monodomain steps/s over 8 ellipsoid shells in one context',
        'value': out['mixed']['steps_per_s'], 'unit': 'steps/s', 'higher_is_better': True, 'n_gpus': 1,
        'dtype': 'f32 (f64 V, dots, residual)', 'data': 'synthetic (generated shells, default cell state)',
        'config': {'workload': '8 truncated prolate-ellipsoid shells (2.5/1.0 -> 3.5/2.0 cm), tetrahedra at h '
                               '%g..%g cm, fibres -60 -> +60 deg, sigma 1.75/0.19, TT06-epi RL dt 0.01 ms, '
                               'apical stimulus -100000 uA/cm^3 1 ms, Jacobi-CG rtol 1e-8; mixed to %g ms'
                               % (h_max, h_min, total * DT), 'n_nodes': n, 'n_elems': n_elems, 'shell_nodes': sizes},
        'paths': out,
        'mixed_over_fp64': round(out['mixed']['common_window']['steps_per_s'] /
                                 out['fp64']['common_window']['steps_per_s'], 3),
        'mixed_over_fp64_window_ms': [warm * DT, first * DT],
        'accuracy_mixed_vs_fp64': {
            'after_steps': first, 'max_dV_mV': float(dv.max()), 'rms_dV_mV': float(np.sqrt((dv ** 2).mean())),
            'activated_nodes': int(both.sum()),
            'activation_mismatched_nodes': int(((a64 >= 0) != (a32 >= 0)).sum()),
            'activation_max_err_ms': float(np.abs(a32[both] - a64[both]).max()) if both.any() else None},
        'note': 'activation tracking is on in the timed steps',
        'regional_adjoint': adj,
    }
    print(json.dumps(line))


def shell_adjoint(h):
    """The batched-tissue workload's gradient: on the coarsest shell (h), TT06
    RL in FP64 (Jacobi-PCG rtol 1e-12), the gradient of the squared voltage
    mismatch at 64 seeded nodes every 0.1 ms over 20 ms with respect to
    regional scalings of g_Na and g_Kr (four regions: quadrants of azimuth),
    by the checkpointed adjoint (checkpoints every 100 steps), against
    Richardson-combined central differences of the device forward run."""
    from chaste_amd import cmx
    from chaste_amd.shells import ellipsoid_shell
    coords, elems, fibres, apex = ellipsoid_shell(h)
    n = coords.shape[0]
    steps, every = 2000, 10
    progress('shells: adjoint setup (%d nodes)' % n)
    sig = cmx.conductivity_tensors(fibres, (1.75, 0.19), axisymmetric=True)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                       stimuli=[(-100000.0, 0.0, 1.0)], stim_index=np.where(apex, 0, -1).astype(np.int32),
                       precision=cmx.FP64, ksp_rtol=1e-12, ksp_atol=0.0, ksp_max_its=2000, sigma_elem=sig,
                       use_graph=USE_GRAPH)
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(cmx.TT06_EPI)])
    names = [nm.split('__')[-1] for nm, _ in cmx.param_info(cmx.TT06_EPI)]
    ids = (names.index('g_Na'), names.index('g_Kr'))
    region = ((np.arctan2(coords[:, 1], coords[:, 0]) + np.pi) // (np.pi / 2)).astype(int).clip(0, 3)
    rng = np.random.default_rng(5)
    obs = np.sort(rng.choice(n, 64, replace=False)).astype(np.int32)

    def params(scal):                                # scal [2, 4]: per parameter and region
        p = np.repeat(defaults[:, None], n, axis=1)
        for k, q in enumerate(ids):
            p[q] *= scal[k][region]
        return p

    def observe(scal):
        m.set_state(y0)
        m.set_params(params(scal))
        out = []
        for j in range(steps // every):
            m.step(j * every * DT, every)
            out.append(m.voltage()[obs])
        return np.array(out)

    data = observe(np.array([[0.7, 1.0, 1.3, 1.0], [1.0, 1.4, 0.8, 1.0]]))
    s1 = np.ones((2, 4))
    m.set_state(y0)
    m.set_params(params(s1))
    t = time.perf_counter()
    m.step(0.0, steps)
    sync()
    t_fwd = time.perf_counter() - t
    progress('shells: adjoint gradient')
    m.set_state(y0)
    t = time.perf_counter()
    loss, gp, _ = m.adjoint(0.0, steps, obs, every, data, 100)
    t_adj = time.perf_counter() - t
    g = np.array([[(gp[q] * defaults[q])[region == r].sum() for r in range(4)] for q in ids])

    def central(d, eps):
        return (((observe(s1 + eps * d) - data) ** 2).sum() - ((observe(s1 - eps * d) - data) ** 2).sum()) / (2 * eps)

    rel = []
    for k in range(2):
        for r in range(4):
            d = np.zeros((2, 4))
            d[k, r] = 1.0
            for eps in (1e-4, 1e-6):                 # as in the adjoint workload: re-difference across a jump
                d1, d2 = central(d, eps), central(d, eps / 2)
                if abs(d1 - d2) <= 1e-5 * abs(d2):
                    break
            fd = (4 * d2 - d1) / 3
            rel.append(abs(g[k, r] - fd) / max(abs(fd), 1e-3 * np.abs(g).max()))
            progress('shells: adjoint %s region %d: %.6e vs %.6e (eps %g)' % (names[ids[k]], r, g[k, r], fd, eps))
    m.close()
    return {'n_nodes': n, 'steps': steps, 'forward_s': round(t_fwd, 4), 'gradient_s': round(t_adj, 4),
            'gradient_over_forward': round(t_adj / t_fwd, 2), 'loss': loss,
            'regional_gradient': {names[q]: [float(x) for x in g[k]] for k, q in enumerate(ids)},
            'fd_max_rel_err': float(max(rel)), 'fd_within_1e-4': int(sum(r <= 1e-4 for r in rel)),
            'fd_checks': len(rel)}


def shell_union_adjoint(args):
    """The batched-tissue gradient on the whole union: the 8 shells of the
    --shells workload in one FP64 context (TT06 RL, Jacobi-PCG rtol 1e-12),
    the squared voltage mismatch at 64 seeded depolarised nodes every 0.1 ms
    over the first --union-steps steps against a run with seeded regional
    scalings, with respect to regional scalings of g_Na and g_Kr in 32
    regions (8 shells x 4 azimuthal quadrants): 64 parameters.  Checkpoints
    every 250 steps; the memory plan nests a second level when one level does
    not fit.  Checked against Richardson-combined central differences of the
    device forward run along seeded random directions."""
    from chaste_amd import cmx
    from chaste_amd.shells import disjoint_union, ellipsoid_shell
    h_max, h_min = args.shell_h
    progress('union adjoint: meshing')
    meshes = [ellipsoid_shell(h) for h in np.linspace(h_max, h_min, 8)]
    sizes = [int(mm[0].shape[0]) for mm in meshes]
    coords, elems, fibres, apex, _ = disjoint_union(meshes)
    del meshes
    sig = cmx.conductivity_tensors(fibres, (1.75, 0.19), axisymmetric=True)
    del fibres
    n = coords.shape[0]
    shell = np.repeat(np.arange(8), sizes)
    quadrant = ((np.arctan2(coords[:, 1], coords[:, 0]) + np.pi) // (np.pi / 2)).astype(int).clip(0, 3)
    region = shell * 4 + quadrant
    steps, every, ck = args.union_steps, 10, 250
    progress('union adjoint: setup (%d nodes, %d tetrahedra)' % (n, elems.shape[0]))
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                       stimuli=[(-100000.0, 0.0, 1.0)], stim_index=np.where(apex, 0, -1).astype(np.int32),
                       precision=cmx.FP64, ksp_rtol=1e-12, ksp_atol=0.0, ksp_max_its=2000, sigma_elem=sig,
                       use_graph=USE_GRAPH)
    n_elems = int(elems.shape[0])
    del coords, elems, sig
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(cmx.TT06_EPI)])
    names = [nm.split('__')[-1] for nm, _ in cmx.param_info(cmx.TT06_EPI)]
    ids = (names.index('g_Na'), names.index('g_Kr'))
    rng = np.random.default_rng(8)
    truth = np.exp(rng.uniform(np.log(0.7), np.log(1.4), (2, 32)))

    def params(scal):                                # scal [2, 32]: per parameter and region
        p = np.repeat(defaults[:, None], n, axis=1)
        for k, q in enumerate(ids):
            p[q] *= scal[k][region]
        return p

    # observation nodes: 64 seeded nodes among those depolarised by the end
    m.set_params(params(truth))
    m.step(0.0, steps)
    V_end = m.voltage()
    cand = np.where(V_end > -60.0)[0]
    obs = np.sort(rng.choice(cand, min(64, len(cand)), replace=False)).astype(np.int32)

    def observe(scal):
        m.set_state(y0)
        m.set_params(params(scal))
        out = []
        for j in range(steps // every):
            m.step(j * every * DT, every)
            out.append(m.voltage()[obs])
        return np.array(out)

    progress('union adjoint: data run')
    data = observe(truth)
    s1 = np.ones((2, 32))
    m.set_state(y0)
    m.set_params(params(s1))
    sync()
    t = time.perf_counter()
    m.step(0.0, steps)
    sync()
    t_fwd = time.perf_counter() - t
    progress('union adjoint: gradient (%d steps, forward %.1f s)' % (steps, t_fwd))
    os.environ['CMX_PROGRESS'] = '1'       # the adjoint reports each replayed piece on stderr
    m.set_state(y0)
    t = time.perf_counter()
    loss, gp, _ = m.adjoint(0.0, steps, obs, every, data, ck)
    t_adj = time.perf_counter() - t
    del os.environ['CMX_PROGRESS']
    st = m.stats()
    g = np.array([np.bincount(region, weights=gp[q] * defaults[q], minlength=32) for q in ids])

    def loss_at(scal):
        return ((observe(scal) - data) ** 2).sum()

    rel = []
    for k in range(args.union_directions):
        d = rng.standard_normal((2, 32))
        d /= np.linalg.norm(d)
        for eps in (1e-4, 1e-6):
            d1 = (loss_at(s1 + eps * d) - loss_at(s1 - eps * d)) / (2 * eps)
            progress('union adjoint: direction %d, central difference at eps %g: %.8e' % (k, eps, d1))
            d2 = (loss_at(s1 + eps / 2 * d) - loss_at(s1 - eps / 2 * d)) / eps
            progress('union adjoint: direction %d, central difference at eps %g: %.8e' % (k, eps / 2, d2))
            if abs(d1 - d2) <= 1e-5 * abs(d2):
                break
        fd = (4 * d2 - d1) / 3
        ad = float((g * d).sum())
        rel.append(abs(ad - fd) / abs(fd))
        progress('union adjoint: direction %d: adjoint %.8e, central difference %.8e (rel %.2e)' % (k, ad, fd, rel[-1]))
    m.close()
    slot_gb = st['n_states'] * ((n + 63) // 64 * 64) * 8 / 1e9
    line = {
        'metric': 'batched tissue gradient: regional conductance-scaling gradient by checkpointed adjoint',
        'value': round(t_adj, 3), 'unit': 's', 'higher_is_better': False, 'n_gpus': 1, 'dtype': 'f64',
        'data': 'synthetic (generated shells, seeded regional scalings)',
        'config': {'workload': '8 truncated prolate-ellipsoid shells in one context (h %g..%g cm), TT06-epi RL dt '
                               '0.01 ms, FP64 Jacobi-PCG rtol 1e-12, %d steps, 64 observation nodes every %d steps, '
                               'g_Na and g_Kr scalings in 32 regions (shell x azimuthal quadrant)'
                               % (h_max, h_min, steps, every), 'n_nodes': n, 'n_elems': n_elems,
                   'shell_nodes': sizes},
        'forward_s': round(t_fwd, 3), 'gradient_over_forward': round(t_adj / t_fwd, 2), 'loss': loss,
        'checkpoints': {'every': ck, 'stored_states': st['adjoint_slots'],
                        'segment_checkpoint_every': st['adjoint_sub_interval'],
                        'state_gb': round(slot_gb, 3), 'stored_gb': round(st['adjoint_slots'] * slot_gb, 1),
                        'one_level_would_store_gb': round(((steps + ck - 1) // ck + ck + 1) * slot_gb, 1)},
        'fd_rel_err': [float(r) for r in rel], 'fd_within_1e-4': int(sum(r <= 1e-4 for r in rel)),
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=4980)   # with the warmup: the 50 ms run
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--roof-steps', type=int, default=50)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-oracle-check', action='store_true',
                    help='skip the oracle comparison of the first %d steps of each accuracy run' % ORACLE_STEPS)
    ap.add_argument('--skip-cfg1', action='store_true', help='skip the LR91 101x101 square (launch-bound case)')
    ap.add_argument('--no-accuracy', action='store_true',
                    help='skip the 50 ms accuracy runs (for runs under a profiler)')
    ap.add_argument('--skip-propagating', action='store_true', help='only the named-stimulus case')
    ap.add_argument('--no-graph', action='store_true', help='plain launches instead of hipGraph replay')
    ap.add_argument('--no-autotune', action='store_true', help='the default solver block count (256)')
    ap.add_argument('--lookup-tables', action='store_true',
                    help='batched cells: gate steady states and decay factors from a table in V')
    ap.add_argument('--no-tables', action='store_true',
                    help='slab: skip the third path (mixed with gate lookup tables)')
    ap.add_argument('--no-kernel-events', action='store_true',
                    help='no per-kernel event timing (ranking / roofline passes); for runs under a profiler')
    ap.add_argument('--adjoint', action='store_true',
                    help='the adjoint workload instead (gradient time and finite-difference check)')
    ap.add_argument('--adjoint-directions', type=int, default=32)
    ap.add_argument('--cells', type=int, default=0,
                    help='the batched-cell workload with this many cells instead (65536 for the named sweep)')
    ap.add_argument('--cells-check', type=int, default=256, help='cells also run in FP64 for the accuracy check')
    ap.add_argument('--cells-model', choices=['ord', 'tt06'], default='ord', help='cell model of the batched-cell sweep')
    ap.add_argument('--lanes-per-cell', type=int, default=0, choices=[0, 1, 2],
                    help='batched cells: 1 or 2 wavefronts per 64 cells; 0 picks per launch by occupancy')
    ap.add_argument('--shells', action='store_true', help='the batched ellipsoid-shell tissue workload instead')
    ap.add_argument('--shell-steps', type=int, default=10000)
    ap.add_argument('--shell-h', type=float, nargs=2, default=(0.04, 0.02), metavar=('H_MAX', 'H_MIN'),
                    help='element sizes of the first and last shell (cm)')
    ap.add_argument('--shell-union-adjoint', action='store_true',
                    help='the regional gradient on the whole 8-shell union instead')
    ap.add_argument('--union-steps', type=int, default=500)
    ap.add_argument('--union-directions', type=int, default=2)
    args = ap.parse_args()
    global USE_GRAPH, AUTOTUNE, LOOKUP_TABLES
    USE_GRAPH = not args.no_graph
    AUTOTUNE = not args.no_autotune
    LOOKUP_TABLES = args.lookup_tables
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU only: the monodomain slab runs on a single MI355X here '
              '(multi-GPU sharding is a design note, not built); run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        adjoint_case(args)
        return
    if args.cells:
        cells_case(args)
        return
    if args.shells:
        shells_case(args)
        return
    if args.shell_union_adjoint:
        shell_union_adjoint(args)
        return
    from chaste_amd import cmx
    coords, elems = cmx.slab_mesh(3, EX, EY, EZ, H)
    n = coords.shape[0]
    stim_index = np.where(coords[:, 0] < STIM_XLIM, 0, -1).astype(np.int32)
    n_states = cmx.model_info(cmx.TT06_EPI)[0]
    y0 = cmx.initial_state(cmx.TT06_EPI)
    init = np.repeat(np.asarray(y0, dtype=np.float64)[:, None], n, axis=1)

    def run_case(stim_mag, with_roofline, oracle_v):
        """Both paths: timing (and ranking / roofline passes), then the
        accuracy run on the same context reset to the initial state: the
        whole 50 ms (ACC_STEPS steps from t = 0, whatever --steps is) with
        activation tracking, the voltage after ORACLE_STEPS steps compared
        with the oracle's FP64 run."""
        results, runs = {}, {}
        paths = [('fp64', cmx.FP64, False), ('mixed', cmx.MIXED, False)]
        if not args.no_tables:
            paths.append(('mixed_tables', cmx.MIXED, True))
        for name, prec, tables in paths:
            progress('%s stim %g: setup' % (name, stim_mag))
            m = build_problem(prec, coords, elems, stim_index, cmx, stim_mag, tables)
            r = time_path(m, args.steps, args.warmup, args.roof_steps if (with_roofline and not tables) else 5,
                          events=not args.no_kernel_events)
            refined = prec == cmx.MIXED
            st = m.stats()
            kinds = {}
            for k, (ms, cnt) in r['ranking'].items():
                work = working_launches_per_step(k, r['rank_its'], refined, cnt / 5) * 5
                b = bytes_per_row(k, prec, n_states, refined) * n
                kinds[k] = {'ms_per_step': ms / 5, 'launches_per_step': cnt / 5,
                            'working_launches_per_step': work / 5,
                            'achieved_GBs': (work * b / (ms * 1e-3) / 1e9) if ms > 0 else None}
            dom = r['dominant']
            dom_launches = max(r['dominant_launches'], 1)
            dom_work = working_launches_per_step(dom, r['roof_its'], refined, dom_launches / args.roof_steps) \
                * args.roof_steps
            dom_bytes = dom_work * bytes_per_row(dom, prec, n_states, refined) * n
            sb = survey_bytes(n, st['nnz'], n_states, prec, r['its'], r['refine'])
            results[name] = {
                'steps_per_s': args.steps / r['elapsed'],
                'ms_per_step': 1e3 * r['elapsed'] / args.steps,
                'cg_its_per_step': r['its'],
                'refine_sweeps_per_step': r['refine'],
                'graphs': r['graphs'],
                'survey_bytes_per_step': sb['step'],
                'survey_step_frac_of_hbm_peak': round(sb['step'] / (r['elapsed'] / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                'kernels_profiled': kinds,
                'dominant_kernel': dom,
                'dominant_avg_launch_ms': r['dominant_ms'] / dom_launches,
                'dominant_bytes_per_launch': dom_bytes / dom_launches,
                'dominant_survey_bytes_per_launch': sb['cell'] if dom == 'cell' else None,
                'format': FORMAT_NAMES[st['format']],
                'n_patterns': st['n_patterns'],
                'autotune': r['autotune'],
            }
            if args.no_accuracy:
                m.close()
                continue
            # accuracy: the same simulation from t = 0 with activation tracking (0 mV)
            progress('%s stim %g: accuracy run (%d steps)' % (name, stim_mag, ACC_STEPS))
            m.set_state(init)
            m.set_activation(0.0, 1, 0.0)
            m.step(0.0, ORACLE_STEPS)
            v_early = m.voltage()
            m.step(ORACLE_STEPS * DT, ACC_STEPS - ORACLE_STEPS)
            runs[name] = (m.voltage(), m.activation()[4], v_early)
            m.close()
        if args.no_accuracy:
            return results, None

        def against_fp64(name):
            dv = np.abs(runs[name][0] - runs['fp64'][0])
            a64, a32 = runs['fp64'][1], runs[name][1]
            both = (a64 >= 0) & (a32 >= 0)
            return {'max_dV_mV': float(dv.max()), 'rms_dV_mV': float(np.sqrt((dv ** 2).mean())),
                    'activation_max_err_ms': float(np.abs(a32[both] - a64[both]).max()) if both.any() else None,
                    'activation_rms_err_ms': float(np.sqrt(((a32[both] - a64[both]) ** 2).mean()))
                    if both.any() else None,
                    'activated_nodes': int(both.sum()),
                    'activation_mismatched_nodes': int(((a64 >= 0) != (a32 >= 0)).sum()),
                    'after_steps': ACC_STEPS}
        acc = against_fp64('mixed')
        acc['run'] = ('the whole 50 ms from t = 0 on each path with activation tracking (first upward crossing '
                      'of 0 mV, interpolated between steps); max/RMS dV over all nodes at 50 ms')
        if 'mixed_tables' in runs:
            acc['mixed_tables'] = against_fp64('mixed_tables')
        if oracle_v is not None:
            for name in ('fp64', 'mixed'):
                d = np.abs(runs[name][2] - oracle_v)
                acc['vs_oracle_%s' % name] = {'after_steps': ORACLE_STEPS, 'max_dV_mV': float(d.max()),
                                              'rms_dV_mV': float(np.sqrt((d ** 2).mean())),
                                              'max_rel': float(d.max() / np.abs(oracle_v).max())}
            acc['vs_oracle_basis'] = ('both runs stop CG at rtol 1e-8 from different initial guesses (oracle: '
                                      'V^n; device: extrapolated), so they agree to the solver tolerance, not to '
                                      '1e-9; the 1e-9 parity runs at rtol 1e-12 from V^n (tests/test_slab_full.py)')
        return results, acc

    cpu_threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    cpu_threads = max(1, min(cpu_threads, os.cpu_count() or 1))
    oracle_named = oracle_prop = None
    if not args.no_oracle_check and not args.no_accuracy:
        progress('oracle: %d FP64 steps of each stimulus on %d threads' % (ORACLE_STEPS, cpu_threads))
        oracle_named = oracle_slab_run(STIM[0], ORACLE_STEPS, cpu_threads)
        if not args.skip_propagating:
            oracle_prop = oracle_slab_run(STIM_PROPAGATING, ORACLE_STEPS, cpu_threads)
    results, accuracy = run_case(STIM[0], True, oracle_named)
    prop_results, prop_accuracy = (None, None) if args.skip_propagating else \
        run_case(STIM_PROPAGATING, False, oracle_prop)
    cfg1 = None if args.skip_cfg1 else cfg1_case(cmx)
    mixed = results['mixed']
    steps_s = mixed['steps_per_s']
    dom = mixed['dominant_kernel']
    dom_ms = mixed['dominant_avg_launch_ms']
    timed_kernels = bool(np.isfinite(dom_ms))
    # roofline.achieved: the survey's algorithmic bytes of one launch of the
    # dominant kernel (cells: n (16 + 2 S s + 4)) over its average duration
    dom_survey = mixed['dominant_survey_bytes_per_launch'] or mixed['dominant_bytes_per_launch']
    achieved = dom_survey / (dom_ms * 1e-3) / 1e9 if timed_kernels else float('nan')
    traffic = None
    tfile = os.path.join(ROOT, 'profiles', 'hbm_traffic_slab442k.json')
    if os.path.exists(tfile):
        t = json.load(open(tfile)).get('mixed', {}).get(dom)
        # bytes moved per launch that does its work (the roofline's launches)
        traffic = round(t['bytes_per_working_dispatch']) if t else None

    line = {
        'metric': 'monodomain steps/s + node-ODE updates/s, FP32-mixed vs FP64 at <=0.5 mV',
        'value': round(steps_s, 2),
        'unit': 'steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(mixed['ms_per_step'], 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 (f64 V, dots, residual)',
        'data': 'synthetic (regular slab, default cell state)',
        'config': {'workload': 'TT06-epi RL monodomain slab 2.0x0.7x0.3 cm h=0.01 (201x71x31 = 442,401 nodes), '
                               'dt 0.01 ms, stim -50000 uA/cm^3 1 ms x<0.03, Jacobi-CG rtol 1e-8'
                               + {0: '', 1: ', initial guess V^n + x_{n-1}',
                                  2: ', initial guess V^n + 2 x_{n-1} - x_{n-2}'}[EXTRAPOLATE]
                               + '; timed: steps %d..%d (t = %g..%g ms) after %d untimed'
                               % (args.warmup + 1, args.warmup + args.steps, args.warmup * DT,
                                  (args.warmup + args.steps) * DT, args.warmup),
                   'n_nodes': n, 'n_elems': int(elems.shape[0])},
        'node_ode_updates_per_s': round(steps_s * n, 1),
        'fp64_steps_per_s': round(results['fp64']['steps_per_s'], 2),
        'fp64_node_ode_updates_per_s': round(results['fp64']['steps_per_s'] * n, 1),
        'mixed_over_fp64': round(steps_s / results['fp64']['steps_per_s'], 3),
        'accuracy_mixed_vs_fp64': accuracy,
        'paths': results,
        'propagating': None if prop_results is None else {
            'note': 'the named stimulus stays below threshold in this tissue (no node activates); the same '
                    'run with %g uA/cm^3 launches a wave across the slab' % STIM_PROPAGATING,
            'stim_uA_per_cm3': STIM_PROPAGATING,
            'steps_per_s': round(prop_results['mixed']['steps_per_s'], 2),
            'fp64_steps_per_s': round(prop_results['fp64']['steps_per_s'], 2),
            'mixed_over_fp64': round(prop_results['mixed']['steps_per_s'] / prop_results['fp64']['steps_per_s'], 3),
            'cg_its_per_step': {k: round(v['cg_its_per_step'], 3) for k, v in prop_results.items()},
            'graphs': {k: v['graphs'] for k, v in prop_results.items()},
            'accuracy_mixed_vs_fp64': prop_accuracy},
        'lookup_tables': None if 'mixed_tables' not in results else {
            'path': 'mixed with each gate\'s steady state and decay factor read from a table in V (-250..350 mV, '
                    '0.02 mV rows, linear interpolation) instead of evaluated (cmx_opts.lookup_tables)',
            'steps_per_s': round(results['mixed_tables']['steps_per_s'], 2),
            'propagating_steps_per_s': None if prop_results is None else
            round(prop_results['mixed_tables']['steps_per_s'], 2),
            'cell_ms_per_step': results['mixed_tables']['kernels_profiled'].get('cell', {}).get('ms_per_step'),
            'cell_ms_per_step_without': results['mixed']['kernels_profiled'].get('cell', {}).get('ms_per_step'),
            'accuracy_vs_fp64': None if accuracy is None else accuracy.get('mixed_tables'),
            'propagating_accuracy_vs_fp64': None if prop_accuracy is None else prop_accuracy.get('mixed_tables')},
        'cfg1_lr91_square': cfg1,
        'roofline': None if not timed_kernels else {
            'bound': 'hbm', 'kernel': dom + ' (mixed)', 'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic,
            'traffic_source': 'profiles/hbm_traffic_slab442k.json (FETCH_SIZE/WRITE_SIZE counters, copy-calibrated)',
            'bytes_per_launch': round(dom_survey),
            'bytes_basis': 'survey algorithmic bytes: n (8 V + 8 Iion + 2 S s + 4 stimulus)',
            'bytes_per_launch_fused': round(mixed['dominant_bytes_per_launch']),
            'fused_basis': 'the kernel\'s own traffic: also settles the previous step (Z, x1, x2 in, V out) and '
                           'forms the extrapolated guess (x0, x_{n-2}, Y, G)',
            'avg_launch_ms': round(dom_ms, 5)},
    }
    if not args.no_cpu_baseline:
        progress('cpu baseline: %d threads, then 1' % cpu_threads)
        cpu_sps, cpu_stats = cpu_baseline(args.cpu_steps, cpu_threads)
        one_sps, one_stats = cpu_baseline(args.cpu_steps, 1)
        line['cpu_baseline'] = {'value': round(cpu_sps, 3), 'unit': 'steps/s', 'cores': cpu_threads, 'kind': 'port',
                                'sample': '%d FP64 steps of the same slab from t = 0 (oracle: Jacobi-PCG, OpenMP over '
                                          'nodes and rows), setup excluded; Ode %.2fs AssRhs %.2fs Ksp %.2fs' %
                                          (cpu_stats['steps'], cpu_stats['t_ode'], cpu_stats['t_rhs'],
                                           cpu_stats['t_ksp']),
                                'cpu_model': cpu_model(), 'host_cpus': os.cpu_count(),
                                'single_thread': {'value': round(one_sps, 3), 'unit': 'steps/s', 'cores': 1,
                                                  'Ode_s': round(one_stats['t_ode'], 3),
                                                  'AssRhs_s': round(one_stats['t_rhs'], 3),
                                                  'Ksp_s': round(one_stats['t_ksp'], 3)},
                                'preconditioner': 'Jacobi (the reference default is bjacobi/ILU(0))'}
    print(json.dumps(line))


if __name__ == '__main__':
    main()
