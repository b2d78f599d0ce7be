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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EX, EY, EZ, H = 200, 70, 30, 0.01
SIGMA = (1.75, 0.19, 0.19)
STIM = (-50000.0, 0.0, 1.0)       # magnitude uA/cm^3, start ms, duration ms
STIM_XLIM = 0.03
DT = 0.01


def build_problem(precision, coords, elems, stim_index, cmx):
    return cmx.Monodomain(coords, elems, np.array(SIGMA), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                          stimuli=[STIM], stim_index=stim_index, precision=precision,
                          dt_pde=DT, dt_ode=DT, ksp_rtol=1e-8, ksp_atol=0.0, ksp_max_its=500,
                          refine_sweeps=1 if precision == cmx.MIXED else 0)


def cell_bytes_per_node(precision, n_states):
    """Algorithmic HBM bytes of one cell-step launch per node: V in (8),
    non-V states in and out, stimulus id (1), forcing F out."""
    s = 4 if precision == 1 else 8
    return 8 + 2 * (n_states - 1) * s + 1 + s


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def time_path(m, steps, warmup, profile_kind):
    """W untimed steps, then exactly K timed steps; events on profile_kind."""
    t = 0.0
    if warmup:
        m.step(t, warmup)
        t += warmup * DT
    # untimed profiling pass over every kernel kind to name the dominant one
    m.set_profiling(['cell', 'residual', 'spmv', 'update', 'pupdate', 'finish'])
    m.kernel_times(reset=True)
    m.step(t, 5)
    t += 5 * DT
    all_times = m.kernel_times(reset=True)
    m.set_profiling([profile_kind])
    st0 = m.stats()
    sync()
    t0 = time.perf_counter()
    m.step(t, steps)
    sync()
    elapsed = time.perf_counter() - t0
    st1 = m.stats()
    ktimes = m.kernel_times(reset=True)
    m.set_profiling([])
    its = (st1['cg_iterations'] - st0['cg_iterations']) / steps
    ref = (st1['refine_sweeps'] - st0['refine_sweeps']) / steps
    return elapsed, its, ref, all_times, ktimes, t + steps * DT


def cpu_baseline(n_steps, threads):
    from oracle import pyoracle
    pyoracle.lib()
    r = pyoracle.mono_run(3, EX, EY, EZ, H, model=pyoracle.TT06_EPI, rush_larsen=True, sigma=SIGMA,
                          stim_mag=STIM[0], stim_dur=STIM[2], stim_start=STIM[1], stim_xlim=STIM_XLIM,
                          t_end=n_steps * DT, ksp_rtol=1e-8, threads=threads)
    s = r['stats']
    loop = s['t_ode'] + s['t_rhs'] + s['t_ksp']
    return s['steps'] / loop, s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=300)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU only: the monodomain slab runs on a single MI355X here '
              '(multi-GPU sharding is a design note, not built); run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    from chaste_amd import cmx
    coords, elems = cmx.slab_mesh(3, EX, EY, EZ, H)
    n = coords.shape[0]
    stim_index = np.where(coords[:, 0] < STIM_XLIM, 0, -1).astype(np.int32)
    n_states = cmx.model_info(cmx.TT06_EPI)[0]

    results = {}
    volts = {}
    for name, prec in (('fp64', cmx.FP64), ('mixed', cmx.MIXED)):
        m = build_problem(prec, coords, elems, stim_index, cmx)
        elapsed, its, ref, all_t, kt, t_end = time_path(m, args.steps, args.warmup, 'cell')
        dominant = max(all_t, key=lambda k: all_t[k][0])
        cell_ms, cell_n = kt['cell']
        st = m.stats()
        results[name] = {
            'steps_per_s': args.steps / elapsed,
            'ms_per_step': 1e3 * elapsed / args.steps,
            'cg_its_per_step': its,
            'refine_sweeps_per_step': ref,
            'kernel_ms_per_step_profiled': {k: v[0] / 5 for k, v in all_t.items()},
            'dominant_kernel': dominant,
            'cell_avg_ms': cell_ms / max(cell_n, 1),
            'format': 'stencil' if st['format'] == cmx.FMT_STENCIL else 'ell',
            'n_patterns': st['n_patterns'],
        }
        volts[name] = m.voltage()
        m.close()

    dv = np.abs(volts['mixed'] - volts['fp64'])
    mixed = results['mixed']
    steps_s = mixed['steps_per_s']
    cell_b = cell_bytes_per_node(1, n_states) * n
    achieved = cell_b / (mixed['cell_avg_ms'] * 1e-3) / 1e9

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
                               'dt 0.01 ms, stim -50000 uA/cm^3 1 ms x<0.03, CG rtol 1e-8',
                   'n_nodes': n, 'n_elems': int(elems.shape[0])},
        'node_ode_updates_per_s': round(steps_s * n, 1),
        'fp64_steps_per_s': round(results['fp64']['steps_per_s'], 2),
        'fp64_node_ode_updates_per_s': round(results['fp64']['steps_per_s'] * n, 1),
        'mixed_over_fp64': round(steps_s / results['fp64']['steps_per_s'], 3),
        'accuracy_mixed_vs_fp64': {'max_dV_mV': float(dv.max()), 'rms_dV_mV': float(np.sqrt((dv ** 2).mean())),
                                   'after_steps': args.warmup + 5 + args.steps},
        'paths': results,
        'roofline': {'bound': 'hbm', 'kernel': 'cell_step (mixed)', 'achieved': round(achieved, 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': None, 'bytes_per_launch': cell_b},
    }
    if not args.no_cpu_baseline:
        threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1))
        cpu_sps, cpu_stats = cpu_baseline(args.cpu_steps, threads)
        line['cpu_baseline'] = {'value': round(cpu_sps, 3), 'unit': 'steps/s', 'cores': threads, 'kind': 'port',
                                'sample': '%d FP64 steps of the same slab (oracle: Jacobi-PCG, OpenMP), setup '
                                          'excluded; Ode %.2fs AssRhs %.2fs Ksp %.2fs' %
                                          (cpu_stats['steps'], cpu_stats['t_ode'], cpu_stats['t_rhs'],
                                           cpu_stats['t_ksp'])}
    print(json.dumps(line))


if __name__ == '__main__':
    main()
