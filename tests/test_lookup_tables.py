"""Gate lookup tables of the mixed-precision Rush-Larsen cells.

The table holds, for every gate and V on a 0.02 mV grid, the steady state and
the decay factor exp(-dt/tau) of one Rush-Larsen step (the reference's *Opt
cells tabulate V-only expressions the same way,
heart/src/odes/AbstractLookupTableCollection.hpp:49).  On the CPU each row is
checked against the oracle's own Rush-Larsen step with V clamped
(orc_cell_compute_except_voltage; pinned against forward Euler in
test_oracle_golden.py): one step is affine in the gate's value, so stepping
from 0 and from 1 gives the decay factor E = y(1) - y(0) and the steady state
y(0) / (1 - E).  Between rows the linear interpolation stays within its
second-derivative bound.  On the GPU the tabulated paths are held to the
mixed-precision bar against the FP64 oracle (0.5 mV, 0.1 ms).
"""
import numpy as np
import pytest

from tests.test_gpu_parity import MIXED_ACT_MS, MIXED_MV, first_crossing

# gate states (Rush-Larsen classification) of each model, in state order
GATES = {
    'lr91': [1, 2, 3, 4, 5, 6],                                  # m h j d f X
    'tt06': [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12],                 # Xr1 Xr2 Xs m h j d f f2 s r
    'ord': list(range(9, 30)) + list(range(31, 38)),             # m .. jca, ffp .. xk1
}


def _models():
    from chaste_amd import cmx
    from oracle import pyoracle
    return [('lr91', cmx.LR91, pyoracle.LR91_CELLML, 0.01), ('tt06', cmx.TT06_EPI, pyoracle.TT06_EPI, 0.01),
            ('ord', cmx.ORD_ENDO, pyoracle.ORD_ENDO, 0.005)]


def _oracle_coefficients(oracle, omodel, gates, V, dt):
    """(steady state, decay factor) of every gate from one oracle RL step at V."""
    y = oracle.model_initial_state(omodel)
    y[0] = V
    y0, y1 = y.copy(), y.copy()
    y0[gates], y1[gates] = 0.0, 1.0
    oracle.cell_compute_except_voltage(omodel, y0, 0.0, dt, dt, integrator=oracle.RL)
    oracle.cell_compute_except_voltage(omodel, y1, 0.0, dt, dt, integrator=oracle.RL)
    E = y1[gates] - y0[gates]
    return y0[gates] / (1.0 - E), E


@pytest.mark.parametrize('name', ['lr91', 'tt06', 'ord'])
def test_gate_table_rows_match_oracle_rush_larsen_step(oracle, name):
    from chaste_amd import cmx
    _, model, omodel, dt = [m for m in _models() if m[0] == name][0]
    gates = GATES[name]
    table, vmin, dv = cmx.gate_table(model, dt)
    assert vmin == -250.0 and dv == 0.02 and table.shape[0] == 30001
    assert table.shape[1] == (2 * len(gates) + 3) // 4 * 4
    assert np.isfinite(table).all()
    for V in (-120.0, -88.0, -47.14, -20.0, 0.0, 14.98, 40.0, 153.6):
        r = int(round((V - vmin) / dv))
        ginf, E = _oracle_coefficients(oracle, omodel, gates, vmin + r * dv, dt)
        np.testing.assert_allclose(table[r, 0:2 * len(gates):2], ginf, rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(table[r, 1:2 * len(gates):2], E, rtol=2e-6, atol=1e-7)
    # between rows: linear interpolation of the neighbouring rows, within
    # h^2/8 max|f''| (f'' from second differences of the rows)
    rng = np.random.default_rng(5)
    for V in rng.uniform(-100.0, 60.0, 12):
        x = (V - vmin) / dv
        r = int(np.floor(x))
        f = x - r
        lerp = table[r] + f * (table[r + 1] - table[r])
        ginf, E = _oracle_coefficients(oracle, omodel, gates, V, dt)
        d2 = np.abs(table[r + 1] - 2 * table[r] + table[r - 1]).astype(np.float64) + np.abs(
            table[r + 2] - 2 * table[r + 1] + table[r]).astype(np.float64)
        bound = d2 / 8 + 3e-7
        assert np.all(np.abs(lerp[0:2 * len(gates):2] - ginf) <= bound[0:2 * len(gates):2] + 2e-6 * np.abs(ginf))
        assert np.all(np.abs(lerp[1:2 * len(gates):2] - E) <= bound[1:2 * len(gates):2] + 2e-6 * np.abs(E))


def test_gate_table_depends_on_step_length():
    """The decay factor is exp(-dt/tau): halving dt takes its square root."""
    from chaste_amd import cmx
    t1, _, _ = cmx.gate_table(cmx.TT06_EPI, 0.01)
    t2, _, _ = cmx.gate_table(cmx.TT06_EPI, 0.005)
    np.testing.assert_allclose(t1[:, 0:22:2], t2[:, 0:22:2], rtol=0, atol=0)
    np.testing.assert_allclose(np.sqrt(t1[:, 1:22:2].astype(np.float64)), t2[:, 1:22:2], rtol=2e-6, atol=1e-7)


@pytest.mark.gpu
def test_tables_tissue_tt06_matches_oracle(oracle):
    """TT06 RL slab with a propagating wave, mixed precision with tables:
    0.5 mV / 0.1 ms of the FP64 oracle at every step, and close to the
    evaluated mixed run."""
    from chaste_amd import cmx
    print('case tables tissue tt06', flush=True)
    dims = (3, 30, 6, 4)
    coords, elems = cmx.slab_mesh(*dims, 0.01)
    stim_index = np.where(coords[:, 0] < 0.03, 0, -1).astype(np.int32)
    traces = {}
    for tables in (False, True):
        m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                           stimuli=[(-100000.0, 0.0, 1.0)], stim_index=stim_index, precision=cmx.MIXED,
                           ksp_rtol=1e-8, refine_sweeps=1, lookup_tables=tables, use_graph=tables)
        steps = 800
        tr = np.zeros((steps + 1, m.n))
        tr[0] = m.voltage()
        for k in range(0, steps, 50):
            m.step(k * 0.01, 50)
            tr[k + 50] = m.voltage()
        m.close()
        traces[tables] = tr
    n = coords.shape[0]
    ref = oracle.mono_run(*dims, model=oracle.TT06_EPI, rush_larsen=True, sigma=(1.75, 0.19, 0.19),
                          stim_mag=-100000.0, stim_dur=1.0, stim_xlim=0.03, t_end=8.0, ksp_rtol=1e-12,
                          trace_nodes=np.arange(n))
    rtrace = np.vstack([np.full(n, traces[True][0][0]), ref['trace']])
    rows = np.arange(0, 801, 50)
    dv = np.abs(traces[True][rows] - rtrace[rows])
    print('tables: max dV vs oracle', dv.max(), 'vs evaluated mixed', np.abs(traces[True] - traces[False])[rows].max())
    assert dv.max() <= MIXED_MV
    assert np.abs(traces[True] - traces[False])[rows].max() <= 0.25


@pytest.mark.gpu
def test_tables_tissue_activation_matches_oracle(oracle):
    """Activation times of the tabulated mixed run, every step, against the
    FP64 oracle (0.1 ms)."""
    from chaste_amd import cmx
    print('case tables activation', flush=True)
    dims = (3, 30, 6, 4)
    coords, elems = cmx.slab_mesh(*dims, 0.01)
    stim_index = np.where(coords[:, 0] < 0.03, 0, -1).astype(np.int32)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                       stimuli=[(-100000.0, 0.0, 1.0)], stim_index=stim_index, precision=cmx.MIXED,
                       ksp_rtol=1e-8, refine_sweeps=1, lookup_tables=True)
    n = m.n
    trace = np.zeros((801, n))
    trace[0] = m.voltage()
    for k in range(800):
        m.step(k * 0.01, 1)
        trace[k + 1] = m.voltage()
    ref = oracle.mono_run(*dims, model=oracle.TT06_EPI, rush_larsen=True, sigma=(1.75, 0.19, 0.19),
                          stim_mag=-100000.0, stim_dur=1.0, stim_xlim=0.03, t_end=8.0, ksp_rtol=1e-12,
                          trace_nodes=np.arange(n))
    rtrace = np.vstack([np.full(n, trace[0][0]), ref['trace']])
    assert np.abs(trace - rtrace).max() <= MIXED_MV
    a1, a2 = first_crossing(trace, 0.01), first_crossing(rtrace, 0.01)
    assert np.array_equal(a1 < 0, a2 < 0) and (a1 >= 0).sum() > 0
    act = a1 >= 0
    print('tables activation err', np.abs(a1[act] - a2[act]).max())
    assert np.abs(a1[act] - a2[act]).max() <= MIXED_ACT_MS


@pytest.mark.gpu
def test_tables_batched_ord_cells_against_fp64():
    """Batched O'Hara-Rudy cells (Rush-Larsen, dt 0.005) with tables against
    the FP64 batched run: final V within 0.5 mV, the same action-potential
    counts, onsets within 0.1 ms.  The set includes a cell whose scalings drive
    V to +153 mV (inside the table's range)."""
    from chaste_amd import cmx
    print('case tables ord cells', flush=True)
    n = 48
    rng = np.random.default_rng(11)
    defaults = np.array([d for _, d in cmx.param_info(cmx.ORD_ENDO)])
    scal = np.exp(rng.uniform(np.log(0.5), np.log(2.0), (len(defaults), n)))
    scal[:, 0] = [0.766, 1.982, 0.514, 0.77, 1.093, 0.838, 1.708]
    params = defaults[:, None] * scal
    period = np.round(rng.uniform(300.0, 1000.0, n))
    period[0] = 358.0
    t_end = period * 2
    stimuli = [(-80.0, 0.0, 2.0, float(p)) for p in period]
    y32, p32, _ = cmx.cells_run(cmx.ORD_ENDO, cmx.RUSH_LARSEN, t_end, 0.005, stimuli=stimuli, params=params,
                                precision=cmx.MIXED, lookup_tables=True)
    y64, p64, _ = cmx.cells_run(cmx.ORD_ENDO, cmx.RUSH_LARSEN, t_end, 0.005, stimuli=stimuli, params=params,
                                precision=cmx.FP64)
    print('tables ord: max dV', np.abs(y32[0] - y64[0]).max())
    assert np.abs(y32[0] - y64[0]).max() <= MIXED_MV
    assert np.array_equal(p32[0], p64[0])
    assert np.abs(p32[1] - p64[1]).max() <= MIXED_ACT_MS


@pytest.mark.gpu
def test_tables_refuse_voltage_outside_and_bad_options():
    """A voltage outside the table fails the step with a clear error (the
    reference's tabulated cells throw), and tables need mixed precision with
    Rush-Larsen cells."""
    from chaste_amd import cmx
    print('case tables errors', flush=True)
    coords, elems = cmx.slab_mesh(2, 6, 6, 0, 0.01)
    kw = dict(stimuli=[(-25000.0, 0.0, 0.5)], stim_index=np.zeros(coords.shape[0], dtype=np.int32))
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.RUSH_LARSEN, precision=cmx.MIXED,
                       lookup_tables=True, **kw)
    V = m.voltage()
    V[5] = 400.0
    m.set_voltage(V)
    with pytest.raises(cmx.CmxError, match='outside the gate lookup table'):
        m.step(0.0, 1)
    m.close()
    for prec, ig in ((cmx.FP64, cmx.RUSH_LARSEN), (cmx.MIXED, cmx.FORWARD_EULER)):
        with pytest.raises(cmx.CmxError, match='lookup tables need'):
            cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, ig, precision=prec, lookup_tables=True,
                           **kw)
