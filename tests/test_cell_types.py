"""Tissues that mix cell types (a cell factory returning different cell
classes per node, heart/src/problem/cell_factories/AbstractCardiacCellFactory.hpp:100):
the device steps each (model, integrator) group with its own launch of the
cell kernel through a node index list.

* Two groups of the same type are the single-group run, bitwise.
* LR91 forward-Euler cells on the left of a 2-D sheet and TT06 Rush-Larsen
  cells on the right, with a wave crossing the boundary, against the
  oracle's per-node-type run: FP64 within 1e-9 (relative to max |V|) at every
  printing step, mixed within 0.5 mV and 0.1 ms activation; also with
  operator splitting.
* The state planes (NaN past each node's own model) round-trip; Iion before
  the first step matches the oracle's cells; autotune restores every group.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (2, 60, 20, 0)


def sheet():
    from chaste_amd import cmx
    coords, elems = cmx.slab_mesh(*DIMS, 0.01)
    left = coords[:, 0] < 0.3
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    return coords, elems, left, stim_index


def make(cmx, coords, elems, stim_index, cell_types=None, model=None, integrator=None, **kw):
    args = dict(stimuli=[(-400000.0, 0.0, 0.5)], stim_index=stim_index, ksp_rtol=1e-12, ksp_atol=0.0, use_graph=True)
    args.update(kw)
    if cell_types is not None:
        return cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.FORWARD_EULER,
                              cell_types=cell_types, **args)
    return cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), model, integrator, **args)


@pytest.mark.parametrize('precision', [0, 1], ids=['fp64', 'mixed'])
def test_two_groups_of_one_type_equal_one_group(precision):
    from chaste_amd import cmx
    print('case cell types: split groups', precision, flush=True)
    coords, elems, left, stim_index = sheet()
    group_of = (np.arange(coords.shape[0]) % 3 == 0).astype(np.int32)     # interleaved
    kw = dict(precision=precision, refine_sweeps=1 if precision else 0, extrapolate=2 if precision else 0)
    a = make(cmx, coords, elems, stim_index, model=cmx.TT06_EPI, integrator=cmx.RUSH_LARSEN, **kw)
    b = make(cmx, coords, elems, stim_index,
             cell_types=(group_of, [(cmx.TT06_EPI, cmx.RUSH_LARSEN), (cmx.TT06_EPI, cmx.RUSH_LARSEN)]), **kw)
    for m in (a, b):
        m.set_activation(0.0, 1, 0.0)
        m.step(0.0, 150)
        m.step(1.5, 150)
    assert np.array_equal(a.voltage(), b.voltage())
    assert np.array_equal(a.state(), b.state())
    assert np.array_equal(a.activation(), b.activation())
    assert b.stats()['graph_launches'] > 0


def _oracle_run(oracle, left, t_end, split=False, trace=None):
    node_model = np.where(left, oracle.LR91_CELLML, oracle.TT06_EPI)
    node_integ = np.where(left, oracle.FE, oracle.RL)
    return oracle.mono_run(*DIMS, sigma=(1.75, 1.75, 1.75), stim_mag=-400000.0, stim_dur=0.5, stim_xlim=0.02,
                           t_end=t_end, ksp_rtol=1e-12, node_model=node_model, node_integrator=node_integ,
                           split=split, trace_nodes=trace, act_threshold=0.0)


@pytest.mark.parametrize('case', ['fp64', 'mixed', 'fp64_split'])
def test_lr91_and_tt06_sheet_matches_oracle(oracle, case):
    from chaste_amd import cmx
    print('case cell types: lr91 | tt06 sheet', case, flush=True)
    coords, elems, left, stim_index = sheet()
    n = coords.shape[0]
    types = (np.where(left, 0, 1).astype(np.int32),
             [(cmx.LR91, cmx.FORWARD_EULER), (cmx.TT06_EPI, cmx.RUSH_LARSEN)])
    mixed = case == 'mixed'
    split = case.endswith('split')
    m = make(cmx, coords, elems, stim_index, cell_types=types, precision=cmx.MIXED if mixed else cmx.FP64,
             refine_sweeps=1 if mixed else 0, ksp_rtol=1e-8 if mixed else 1e-12, operator_splitting=split)
    assert m.stats()['n_states'] == 19
    steps, every = 1200, 50
    m.set_activation(0.0, 1, 0.0)
    printed = []
    for j in range(steps // every):
        m.step(j * every * 0.01, every)
        printed.append(m.voltage())
    ref = _oracle_run(oracle, left, steps * 0.01, split=split, trace=np.arange(n))
    rp = ref['trace'][every - 1::every]
    got = np.array(printed)
    a_gpu, a_ref = m.activation()[4], ref['act']
    both = (a_gpu >= 0) & (a_ref >= 0)
    # the wave has crossed from the LR91 cells into the TT06 cells
    assert both[left].any() and both[~left].sum() > 100
    print('%s: max |dV| %.3e mV, rel %.3e, act err %.3e ms' % (case, np.abs(got - rp).max(),
          np.abs(got - rp).max() / np.abs(rp).max(), np.abs(a_gpu[both] - a_ref[both]).max()))
    one = (a_gpu >= 0) != (a_ref >= 0)
    if mixed:
        assert np.abs(got - rp).max() <= 0.5
        assert np.all(np.maximum(a_gpu, a_ref)[one] >= steps * 0.01 - 0.1)
        assert np.abs(a_gpu[both] - a_ref[both]).max() <= 0.1
    else:
        assert np.abs(got - rp).max() / np.abs(rp).max() <= 1e-9
        assert not one.any()
        assert np.abs(a_gpu[both] - a_ref[both]).max() <= 1e-6


def test_cell_type_state_iionic_and_autotune(oracle):
    from chaste_amd import cmx
    print('case cell types: state, iionic, autotune', flush=True)
    coords, elems, left, stim_index = sheet()
    types = (np.where(left, 0, 1).astype(np.int32),
             [(cmx.LR91, cmx.FORWARD_EULER), (cmx.TT06_EPI, cmx.RUSH_LARSEN)])
    m = make(cmx, coords, elems, stim_index, cell_types=types)
    S = m.state()
    assert S.shape == (19, coords.shape[0])
    y_lr, y_tt = cmx.initial_state(cmx.LR91), cmx.initial_state(cmx.TT06_EPI)
    assert np.array_equal(S[:8, left], np.repeat(np.asarray(y_lr)[:, None], left.sum(), axis=1))
    assert np.isnan(S[8:, left]).all()
    assert np.array_equal(S[:, ~left], np.repeat(np.asarray(y_tt)[:, None], (~left).sum(), axis=1))
    ii = m.iionic()
    assert np.allclose(ii[left], oracle.cell_iionic(oracle.LR91_CELLML, np.asarray(y_lr)), rtol=1e-12, atol=1e-14)
    assert np.allclose(ii[~left], oracle.cell_iionic(oracle.TT06_EPI, np.asarray(y_tt)), rtol=1e-12, atol=1e-14)
    m.step(0.0, 100)
    S1, V1 = m.state(), m.voltage()
    m.autotune(1.0, 10, candidates=[1, 2])
    assert np.array_equal(m.state(), S1, equal_nan=True) and np.array_equal(m.voltage(), V1)
    # a state written back is the state read (NaN padding ignored)
    S1[0] += 0.5
    m.set_state(S1)
    assert np.array_equal(m.state(), S1, equal_nan=True)
    with pytest.raises(cmx.CmxError):
        m.set_params(np.ones((1, coords.shape[0])))
