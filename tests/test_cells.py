"""Batched single cells (cmx_cells_run) against the oracle's single-cell runs.

FP64: final states to 1e-9 of each state's scale and the CellProperties
summary (AP count, last onset, peak, end) to 1e-6 ms / mV; mixed (FP32 states,
FP64 V): final V within 0.5 mV, onsets within 0.1 ms, the ends of the action
potentials (slow repolarisation) within 0.5 ms, the same AP counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sweep(model, n, seed, beats=(1, 3), cl=(300.0, 800.0)):
    """A parameter/pacing sweep: per-cell conductance scalings log-uniform in
    [0.5, 2], cycle lengths uniform in cl, 2 ms stimuli at the start of each
    cycle, uneven run lengths (a whole number of cycles)."""
    from chaste_amd import cmx
    rng = np.random.default_rng(seed)
    defaults = np.array([d for _, d in cmx.param_info(model)])
    scal = np.exp(rng.uniform(np.log(0.5), np.log(2.0), (len(defaults), n)))
    params = defaults[:, None] * scal
    period = np.round(rng.uniform(*cl, n))
    n_beats = rng.integers(beats[0], beats[1] + 1, n)
    mag = -80.0 if model == cmx.TT06_EPI else -25.5
    stimuli = [(mag, 0.0, 2.0, float(p)) for p in period]
    return params, stimuli, period * n_beats


def oracle_cells(oracle, model, integrator, dt, params, stimuli, t_end, threshold):
    omodel = oracle.TT06_EPI if model == 1 else oracle.LR91_CELLML
    S = oracle.num_states(omodel)
    ys, props = [], []
    for i in range(len(t_end)):
        y = oracle.model_initial_state(omodel)
        mag, start, dur, period = stimuli[i][:4]
        v = oracle.cell_trace_v(omodel, y, t_end[i], dt, stim=(mag, dur, start, period), integrator=integrator,
                                params=params[:, i])
        ys.append(y[:S])
        props.append(oracle.cell_properties(v, dt, threshold))
    return np.array(ys).T, np.array(props).T


CASES = {'tt06_rl': (1, 1, 0.005), 'lr91_fe': (0, 0, 0.01)}


@pytest.mark.parametrize('case', sorted(CASES))
def test_cells_fp64_matches_oracle(oracle, case):
    from chaste_amd import cmx
    model, integrator, dt = CASES[case]
    print('case cells fp64', case, flush=True)
    params, stimuli, t_end = sweep(model, 10, seed=1)
    y, props, stats = cmx.cells_run(model, integrator, t_end, dt, stimuli=stimuli, params=params,
                                    precision=cmx.FP64, threshold=-30.0)
    ry, rprops = oracle_cells(oracle, model, integrator, dt, params, stimuli, t_end, -30.0)
    scale = np.abs(ry).max(axis=1, keepdims=True)
    err = (np.abs(y - ry) / scale).max()
    print('state rel err', err, 'props', np.abs(props - rprops).max(), 'stats', stats)
    assert err <= 1e-9
    # (a cell paced once at a cycle shorter than its AP has no completed AP)
    assert np.array_equal(props[0], rprops[0]) and (props[0] >= 1).sum() >= len(t_end) // 2
    assert np.abs(props[1:] - rprops[1:]).max() <= 1e-6
    assert stats['cell_steps'] == int(np.round(t_end / dt).sum())


def test_cells_mixed_matches_oracle(oracle):
    from chaste_amd import cmx
    print('case cells mixed', flush=True)
    params, stimuli, t_end = sweep(1, 10, seed=2)
    y, props, _ = cmx.cells_run(1, 1, t_end, 0.005, stimuli=stimuli, params=params, precision=cmx.MIXED,
                                threshold=-30.0)
    ry, rprops = oracle_cells(oracle, 1, 1, 0.005, params, stimuli, t_end, -30.0)
    dv = np.abs(y[0] - ry[0])
    dt_on, dt_end = np.abs(props[1] - rprops[1]), np.abs(props[3] - rprops[3])
    print('mixed dV', dv.max(), 'onset', dt_on.max(), 'end', dt_end.max(), 'peak', np.abs(props[2] - rprops[2]).max())
    assert np.array_equal(props[0], rprops[0])
    assert dv.max() <= 0.5
    assert dt_on.max() <= 0.1
    # repolarisation crosses -30 mV at ~1 mV/ms, so the 0.5 mV budget on V
    # allows ~0.5 ms there (measured 0.37 ms after a 300 ms plateau in FP32)
    assert dt_end.max() <= 0.5


def test_cells_launch_chunking_is_exact():
    """Cells keep their state in memory between launches at its own precision:
    any steps-per-launch gives bitwise the same answer; uneven lengths are
    handled by the sorted order (the longest cells first)."""
    from chaste_amd import cmx
    print('case cells chunking', flush=True)
    params, stimuli, t_end = sweep(1, 300, seed=3, beats=(1, 2), cl=(300.0, 400.0))
    ref = cmx.cells_run(1, 1, t_end, 0.005, stimuli=stimuli, params=params, precision=cmx.MIXED)
    for chunk in (77, 100000):
        got = cmx.cells_run(1, 1, t_end, 0.005, stimuli=stimuli, params=params, precision=cmx.MIXED,
                            steps_per_launch=chunk)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    # a cell's answer does not depend on its neighbours in the batch
    one = cmx.cells_run(1, 1, t_end[5:6], 0.005, stimuli=stimuli[5:6], params=params[:, 5:6], precision=cmx.MIXED)
    assert np.array_equal(one[0][:, 0], ref[0][:, 5])


def test_cells_reject_bad_input():
    from chaste_amd import cmx
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(1, 1, [1.0025], 0.005)           # not a whole number of steps
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(1, cmx.BACKWARD_EULER, [1.0], 0.005)
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(0, 0, [1.0], 0.01, params=np.ones((4, 1)))
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(1, 1, [10.0], 0.005, stimuli=[(-80.0, 0.0, 5.0, 2.0)])   # duration > period


@pytest.mark.parametrize('model_name', ['lr91', 'tt06', 'ord'])
def test_two_wavefronts_per_cell_match_one_lane(model_name):
    """lanes_per_cell = 2 (one wavefront advancing the gates, one the rest, the
    gates and V exchanged through LDS each step) runs the one-lane step's
    arithmetic split in two: FP64 final states within 1e-12 of each state's
    scale and the same CellProperties summary; mixed within 1e-3 mV.  Uneven
    lengths (cells ending inside a block of 64) and a partial last block are
    covered; the automatic layout splits a sweep this small."""
    from chaste_amd import cmx
    model, integrator, dt = {'lr91': (cmx.LR91, 0, 0.01), 'tt06': (cmx.TT06_EPI, 1, 0.005),
                             'ord': (cmx.ORD_ENDO, 1, 0.005)}[model_name]
    print('case cells two wavefronts', model_name, flush=True)
    params, stimuli, t_end = sweep(model, 150, seed=7, beats=(1, 2), cl=(300.0, 500.0))
    if model == cmx.ORD_ENDO:
        stimuli = [(-80.0, 0.0, 2.0, s[3]) for s in stimuli]
    for prec, tol in ((cmx.FP64, 1e-12), (cmx.MIXED, None)):
        y1, p1, s1 = cmx.cells_run(model, integrator, t_end, dt, stimuli=stimuli, params=params, precision=prec,
                                   lanes_per_cell=1, steps_per_launch=3000)
        y2, p2, s2 = cmx.cells_run(model, integrator, t_end, dt, stimuli=stimuli, params=params, precision=prec,
                                   lanes_per_cell=2, steps_per_launch=3000)
        assert s1['split_launches'] == 0 and s2['split_launches'] == s2['launches'] > 0
        assert np.isfinite(y2).all()
        assert np.array_equal(p1[0], p2[0])
        if tol is not None:
            scale = np.abs(y1).max(axis=1, keepdims=True) + 1e-300
            print('two wavefronts fp64 rel', (np.abs(y2 - y1) / scale).max())
            assert (np.abs(y2 - y1) / scale).max() <= tol
            assert np.abs(p2[1:] - p1[1:]).max() <= 1e-9
        else:
            print('two wavefronts mixed dV', np.abs(y2[0] - y1[0]).max())
            assert np.abs(y2[0] - y1[0]).max() <= 1e-3
    # automatic: 150 cells are 3 wavefronts, far below half the SIMDs, so every launch splits
    ya, pa, sa = cmx.cells_run(model, integrator, t_end, dt, stimuli=stimuli, params=params, precision=cmx.MIXED,
                               steps_per_launch=3000)
    assert sa['split_launches'] == sa['launches'] > 0 and np.array_equal(ya, y2) and np.array_equal(pa, p2)
    with pytest.raises(cmx.CmxError, match='lanes_per_cell'):
        cmx.cells_run(model, integrator, t_end, dt, precision=cmx.MIXED, lanes_per_cell=3)
