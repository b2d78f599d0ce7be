"""Further generated cell models: Hodgkin-Huxley 1952, Fox 2002, Mahajan
2008, and -- in seconds or absolute currents, converted by the generator --
DiFrancesco-Noble 1985, Faber-Rudy 2000, Maleckar 2008 and Noble-Varghese-Kohl-
Noble 1998 (whose calcium equation uses two other derivatives)
(chaste_amd/codegen/models.py, from the reference's CellML files).

The oracle runs the generated FP64 model; it is pinned here against the
reference's ValidData files and protocols (heart/test/ionicmodels/
TestIonicModels.hpp:305-333 HH52, :700-748 Fox 2002, :910-983 Mahajan 2008,
:579-697 Faber-Rudy 2000 with its endo/mid/epi parameter sets, :883-905
DiFrancesco-Noble 1985, :987-1022 Maleckar 2008, :102-142 and :752-790 N98;
voltages within 1e-3 mV at every sample, RunAndCheckIonicModels.hpp:196-221),
the GetIIonic values those tests hard-code, and backward Euler against
forward Euler with the reference's CompareCellModelResults rule
(RunAndCheckIonicModels.hpp:224-280).  Rush-Larsen is pinned as for LR91
(test_oracle_golden.py): equal to forward Euler with V clamped, and first
order against a dt 1e-6 forward-Euler solution.  On the GPU the batched
cells and a propagating slab are compared with the oracle.
"""
import numpy as np
import pytest

from conftest import load_dat
from tests.test_gpu_parity import MIXED_ACT_MS, MIXED_MV

HH52, FOX, MAHAJAN, DN85, FR2000, MALECKAR, N98 = 4, 5, 6, 7, 8, 9, 10        # oracle model ids


def _run(oracle, model, t_end, dt, stim, per, integ=0, y=None):
    y = oracle.model_initial_state(model) if y is None else y
    out = oracle.cell_compute(model, y, 0.0, t_end, dt, stim=stim, steps_per_sample=per, integrator=integ)
    return y, out


def test_hh52_valid_data(oracle):
    _, ref = load_dat('HH52RegResultValidData.dat')
    y, out = _run(oracle, HH52, 150.0, 0.01, (-20.0, 0.5, 10.0, -1.0), 100)
    assert out.shape[0] == ref.shape[0] == 151
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    # the same cell run again from t = 0 for 15 ms (TestIonicModels.hpp:332-333)
    _run(oracle, HH52, 15.0, 0.01, (-20.0, 0.5, 10.0, -1.0), 100, y=y)
    assert abs(oracle.cell_iionic(HH52, y) - 40.6341) < 1e-3


def test_fox2002_valid_data(oracle):
    hdr, ref = load_dat('FoxRegularStimValidData.dat')
    y, out = _run(oracle, FOX, 200.0, 0.002, (-80.0, 1.0, 50.0, 500.0), 500)
    assert out.shape[0] == ref.shape[0] == 201
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    # every state of the file (5 significant figures), in the model's order
    for j in range(1, 13):
        rel = np.abs(out[:, j] - ref[:, j + 1]) / np.maximum(np.abs(ref[:, j + 1]), 1e-12)
        assert rel.max() < 2e-3, hdr[j + 1]


def test_mahajan2008_valid_data(oracle):
    _, ref = load_dat('Mahajan2008ValidData.dat')
    y0 = oracle.model_initial_state(MAHAJAN)
    assert abs(oracle.cell_iionic(MAHAJAN, y0) - 0.0027) < 1e-3
    y, out = _run(oracle, MAHAJAN, 80.0, 0.001, (-1800.0, 0.05, 10.0, 1000.0), 1000)
    assert out.shape[0] == ref.shape[0] == 81
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3


def test_dn85_valid_data(oracle):
    """Self-excitatory Purkinje cell (seconds and nA in the CellML; no
    stimulus), 1800 ms at dt 0.01, samples every 10 ms."""
    _, ref = load_dat('DiFrancescoNobleValidData.dat')
    y, out = _run(oracle, DN85, 1800.0, 0.01, (0.0, 0.0, 0.0, -1.0), 1000)
    assert out.shape[0] == ref.shape[0] == 181
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    assert abs(oracle.cell_iionic(DN85, y) - (-0.0141)) < 1e-3


def _fr2000(oracle, stim_start, params=None):
    y = oracle.model_initial_state(FR2000)
    return oracle.cell_compute(FR2000, y, 0.0, 1000.0, 0.007, stim=(-25.5, 2.0, stim_start, -1.0),
                               steps_per_sample=500, params=params)


@pytest.mark.parametrize('case', ['DelayedStim', 'Endo', 'Mid', 'Epi'])
def test_fr2000_valid_data(oracle, case):
    """Faber-Rudy 2000 (seconds in the CellML) at dt 0.007, samples every
    3.5 ms; the endo/mid/epi cells set ScaleFactorGks, ScaleFactorIto and
    ScaleFactorGkr (parameters 0-2; g_Na stays at its default)."""
    _, ref = load_dat('FR2000%sValidData.dat' % case)
    sets = {'Endo': (0.462, 0.0, 1.0), 'Mid': (1.154, 0.85, 1.0), 'Epi': (1.154, 1.0, 1.0)}
    if case == 'DelayedStim':
        out = _fr2000(oracle, 10.0)
    else:
        out = _fr2000(oracle, 0.0, params=np.array(list(sets[case]) + [16.0]))
    assert out.shape[0] == ref.shape[0] == 286
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3


def test_maleckar2008_valid_data(oracle):
    """Human atrial cell (seconds, pA and nF in the CellML): the stimulus
    -5.6 uA/cm^2 reaches the model as -280 pA through its 50 pF."""
    _, ref = load_dat('Maleckar2009ValidData.dat')
    y, out = _run(oracle, MALECKAR, 25.0, 0.001, (-5.6, 6.0, 4.0, 1000.0), 1000)
    assert out.shape[0] == ref.shape[0] == 26
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    assert abs(oracle.cell_iionic(MALECKAR, y) - 1.4426) < 1e-3


N98_STIM = (-3.0 / 0.095, 3.0, 10.0, -1.0)


def test_n98_valid_data(oracle):
    """Forward Euler at dt 0.01 for 150 ms, samples every ms; GetIIonic 0.2462."""
    _, ref = load_dat('N98RegResultValidData.dat')
    y, out = _run(oracle, N98, 150.0, 0.01, N98_STIM, 100)
    assert out.shape[0] == ref.shape[0] == 151
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    assert abs(oracle.cell_iionic(N98, y) - 0.2462) < 1e-3


def test_n98_backward_euler_valid_data(oracle):
    """The backward-Euler cell at dt 0.2, every step sampled, against the
    reference's backward-Euler results; GetIIonic 0.2462."""
    _, ref = load_dat('N98BackwardResultValidData.dat')
    y, out = _run(oracle, N98, 150.0, 0.2, N98_STIM, 1, integ=oracle.BE)
    assert out.shape[0] == ref.shape[0]
    assert np.abs(out[:, 0] - ref[:, 1]).max() < 1e-3
    assert abs(oracle.cell_iionic(N98, y) - 0.2462) < 1e-3


def _compare_cell_results(v1, v2, tol, extra=()):
    """CompareCellModelResults on equal sample times: 25x the tolerance where
    the second trace moved more than 0.05 mV since its previous sample."""
    last = v2[0]
    for i in range(len(v2)):
        t = tol * 25 if abs(v2[i] - last) > 0.05 else tol
        last = v2[i]
        assert abs(v1[i] - v2[i]) <= t, (i, v1[i], v2[i])
        for a, b, scale in extra:
            assert abs(a[i] - b[i]) <= t * scale, (i, a[i], b[i])


def test_fox2002_backward_euler_against_forward(oracle):
    stim = (-80.0, 1.0, 50.0, 500.0)
    _, fe = _run(oracle, FOX, 200.0, 0.002, stim, 500)
    _, be = _run(oracle, FOX, 200.0, 0.01, stim, 100, integ=oracle.BE)
    # calcium (Ca_i, state 11) at tol/100 and the h gate (state 2) at tol/10
    _compare_cell_results(fe[:, 0], be[:, 0], 0.2, extra=((fe[:, 11], be[:, 11], 0.01), (fe[:, 2], be[:, 2], 0.1)))


def test_mahajan2008_backward_euler_against_forward(oracle):
    stim = (-1800.0, 0.05, 5.0, -1.0)
    ybe, be = _run(oracle, MAHAJAN, 50.0, 0.005, stim, 100, integ=oracle.BE)
    assert abs(oracle.cell_iionic(MAHAJAN, ybe) - 0.1244) < 1e-3
    _, fe = _run(oracle, MAHAJAN, 50.0, 0.005, stim, 100)
    _compare_cell_results(fe[:, 0], be[:, 0], 0.03)


ALL = [HH52, FOX, MAHAJAN, DN85, FR2000, MALECKAR, N98]
ALL_IDS = ['hh52', 'fox2002', 'mahajan2008', 'dn85', 'fr2000', 'maleckar2008', 'n98']


def _smooth(oracle, model):
    """Mask of the states whose right-hand side is smooth: Faber-Rudy 2000's
    calcium-release trackers (APtrack*, OVRLDtrack*, Cainfluxtrack) switch
    between branches when a state crosses a threshold, and the switching time
    moves with dt, so they do not converge at first order."""
    from chaste_amd import cmx
    cid = {HH52: cmx.HH52, FOX: cmx.FOX2002, MAHAJAN: cmx.MAHAJAN2008, DN85: cmx.DN85, FR2000: cmx.FR2000,
           MALECKAR: cmx.MALECKAR2008, N98: cmx.N98}[model]
    return np.array(['track' not in n for n in cmx.state_names(cid)])


def _clamped_reference(oracle, model, y0):
    """Forward Euler over 1 ms with V clamped at dt 1e-6, and its error
    estimated from the run at 2e-6 (first order: their difference)."""
    ref, ref2 = y0.copy(), y0.copy()
    oracle.cell_compute_except_voltage(model, ref, 0.0, 1.0, 1e-6, integrator=oracle.FE)
    oracle.cell_compute_except_voltage(model, ref2, 0.0, 1.0, 2e-6, integrator=oracle.FE)
    return ref, np.abs(ref - ref2)


@pytest.mark.parametrize('model', ALL, ids=ALL_IDS)
def test_backward_euler_first_order_clamped(oracle, model):
    """With V clamped, backward Euler over 1 ms converges at first order to
    the dt 1e-6 forward-Euler solution (ratio 2 +- 0.2 between dt 0.002 and
    0.001 for every smooth state whose error is above the reference's)."""
    y0 = oracle.model_initial_state(model)
    y0[0] = -20.0
    ref, ref_err = _clamped_reference(oracle, model, y0)
    errs = []
    for dt in (0.002, 0.001):
        y = y0.copy()
        oracle.cell_compute_except_voltage(model, y, 0.0, 1.0, dt, integrator=oracle.BE)
        errs.append(y - ref)
    # states whose error is well above the reference's own
    moved = (np.abs(errs[1]) > 50.0 * ref_err + 1e-13 * np.maximum(np.abs(y0), 1.0)) & _smooth(oracle, model)
    moved[0] = False
    assert moved.sum() >= 3
    ratio = errs[0][moved] / errs[1][moved]
    assert np.all(np.abs(ratio - 2.0) <= 0.2), ratio


@pytest.mark.parametrize('model', ALL, ids=ALL_IDS)
def test_rush_larsen_against_forward_euler_clamped(oracle, model):
    """ComputeExceptVoltage over 1 ms at dt 1e-5, where forward Euler is
    accurate even for the fastest gate (Faber-Rudy's Ito zdv): Rush-Larsen
    and forward Euler agree to 1e-6 relative in every state
    (TestRushLarsen.hpp:130-138),
    and Rush-Larsen converges at first order to the dt 1e-6 forward-Euler
    solution with V clamped at -20 mV."""
    y0 = oracle.model_initial_state(model)
    a, b = y0.copy(), y0.copy()
    oracle.cell_compute_except_voltage(model, a, 0.0, 1.0, 1e-5, integrator=oracle.RL)
    oracle.cell_compute_except_voltage(model, b, 0.0, 1.0, 1e-5, integrator=oracle.FE)
    assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1.0))
    y0[0] = -20.0
    ref, ref_err = _clamped_reference(oracle, model, y0)
    errs = []
    for dt in (0.002, 0.001):
        y = y0.copy()
        oracle.cell_compute_except_voltage(model, y, 0.0, 1.0, dt, integrator=oracle.RL)
        errs.append(y - ref)
    scale = np.maximum(np.abs(y0), 1.0)
    # a gate with V clamped is linear with constant coefficients, where the
    # Rush-Larsen update is the exact solution at any dt: its error is the
    # reference's own (forward Euler at 1e-6, ~1e-7 relative)
    exact = np.abs(errs[0] - errs[1]) <= 1e-12 * scale
    exact[0] = False
    assert np.all(np.abs(errs[0][exact]) <= 2.0 * ref_err[exact] + 1e-12 * scale[exact])
    moved = (np.abs(errs[1]) > 50.0 * ref_err + 1e-13 * scale) & ~exact & _smooth(oracle, model)
    moved[0] = False
    assert exact.sum() + moved.sum() >= 3
    ratio = errs[0][moved] / errs[1][moved]
    assert np.all(np.abs(ratio - 2.0) <= 0.2), ratio


CASES = {
    # name: (cmx id, oracle id, dt of the single-cell runs (FE, RL), cell stimulus (mag, dur, period))
    'hh52': ('HH52', HH52, (0.01, 0.01), (-20.0, 0.5, 20.0)),
    'fox2002': ('FOX2002', FOX, (0.002, 0.002), (-80.0, 1.0, 300.0)),
    'mahajan2008': ('MAHAJAN2008', MAHAJAN, (0.001, 0.002), (-1800.0, 0.05, 300.0)),
    'dn85': ('DN85', DN85, (0.01, 0.01), (-50.0, 1.0, 400.0)),
    'fr2000': ('FR2000', FR2000, (0.007, 0.007), (-25.5, 2.0, 400.0)),
    'maleckar2008': ('MALECKAR2008', MALECKAR, (0.001, 0.002), (-5.6, 6.0, 400.0)),
    'n98': ('N98', N98, (0.01, 0.01), (-3.0 / 0.095, 3.0, 400.0)),
}


def test_library_model_table_matches_oracle(oracle):
    """Each id of the library's model table carries the generated model the
    oracle runs under the same name (initial state, parameter defaults)."""
    from chaste_amd import cmx
    for name, (cid, oid, _, _) in CASES.items():
        y = np.asarray(cmx.initial_state(getattr(cmx, cid)))
        np.testing.assert_array_equal(y, oracle.model_initial_state(oid), err_msg=name)
        assert len(cmx.param_info(getattr(cmx, cid))) >= 3


@pytest.mark.gpu
@pytest.mark.parametrize('integ', ['fe', 'rl'])
@pytest.mark.parametrize('name', sorted(CASES))
def test_batched_cells_fp64_matches_oracle(oracle, name, integ):
    """cmx_cells_run in FP64 against the oracle's cells: every state within
    1e-9 of its scale, the CellProperties summary within 1e-9 ms."""
    from chaste_amd import cmx
    cid, oid, dts, (mag, dur, per0) = CASES[name]
    model = getattr(cmx, cid)
    print('case more-model cells', name, integ, flush=True)
    rng = np.random.default_rng(5)
    n = 40
    dt = dts[1] if integ == 'rl' else dts[0]
    defaults = np.array([d for _, d in cmx.param_info(model)])
    params = defaults[:, None] * np.exp(rng.uniform(np.log(0.8), np.log(1.25), (len(defaults), n)))
    period = np.round(rng.uniform(per0, 2 * per0, n))
    t_end = np.round(np.minimum(period * 1.5, 400.0) / dt) * dt
    stimuli = [(mag, 0.0, dur, float(p)) for p in period]
    ig = cmx.RUSH_LARSEN if integ == 'rl' else cmx.FORWARD_EULER
    y, props, _ = cmx.cells_run(model, ig, t_end, dt, stimuli=stimuli, params=params, precision=cmx.FP64,
                                threshold=-30.0)
    oi = oracle.RL if integ == 'rl' else oracle.FE
    scale = np.abs(np.asarray(cmx.initial_state(model))) + 1e-3
    worst = 0.0
    for c in range(0, n, 9):
        yc = oracle.model_initial_state(oid)
        v = oracle.cell_trace_v(oid, yc, t_end[c], dt, stim=(mag, dur, 0.0, period[c]), integrator=oi,
                                params=params[:, c])
        worst = max(worst, float((np.abs(y[:, c] - yc) / scale).max()))
        pr = oracle.cell_properties(v, dt, -30.0)
        assert pr[0] == props[0][c] and pr[0] >= 1
        assert np.abs(pr[1:] - props[1:, c]).max() <= 1e-9
    print('%s cells %s worst scaled state error %.3e' % (name, integ, worst))
    assert worst <= 1e-9


TISSUE = {
    # name: (stimulus uA/cm^3, duration ms, dt_ode)
    'hh52': (-100000.0, 0.5, 0.01),
    'fox2002': (-400000.0, 1.0, 0.002),
    'mahajan2008': (-2600000.0, 0.1, 0.002),
    'dn85': (-500000.0, 1.0, 0.01),
    'fr2000': (-500000.0, 2.0, 0.005),
    'maleckar2008': (-300000.0, 2.0, 0.002),
    'n98': (-500000.0, 1.0, 0.01),
}


@pytest.mark.gpu
@pytest.mark.parametrize('precision', [0, 1], ids=['fp64', 'mixed'])
@pytest.mark.parametrize('name', sorted(CASES))
def test_tissue_matches_oracle(oracle, name, precision):
    """A 2-D slab with a wave crossing it (Rush-Larsen cells), 4 ms: FP64
    within 1e-9 relative of the oracle; mixed within 0.5 mV and 0.1 ms in
    activation time of the oracle's FP64 run."""
    from chaste_amd import cmx
    cid, oid, _, _ = CASES[name]
    mag, dur, dt_ode = TISSUE[name]
    print('case more-model tissue', name, precision, flush=True)
    dims = (2, 40, 8, 0)
    coords, elems = cmx.slab_mesh(*dims, 0.01)
    n = coords.shape[0]
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), getattr(cmx, cid), cmx.RUSH_LARSEN,
                       stimuli=[(mag, 0.0, dur)], stim_index=stim_index, precision=precision,
                       dt_pde=0.01, dt_ode=dt_ode, ksp_rtol=1e-12 if precision == 0 else 1e-8, ksp_atol=0.0,
                       refine_sweeps=1 if precision else 0, use_graph=True)
    m.set_activation(0.0, 1, 0.0)
    m.step(0.0, 400)
    V, act = m.voltage(), m.activation()[4]
    m.close()
    ref = oracle.mono_run(*dims, model=oid, rush_larsen=True, sigma=(1.75, 1.75, 1.75), stim_mag=mag, stim_dur=dur,
                          stim_xlim=0.02, dt_ode=dt_ode, t_end=4.0, ksp_rtol=1e-12, act_threshold=0.0)
    both = (act >= 0) & (ref['act'] >= 0)
    print('%s tissue: rel err %.3e, max dV %.3e mV, activated %d of %d, act err %.3e' % (
        name, np.abs(V - ref['V']).max() / np.abs(ref['V']).max(), np.abs(V - ref['V']).max(), both.sum(), n,
        np.abs(act[both] - ref['act'][both]).max() if both.any() else -1))
    assert np.array_equal(act >= 0, ref['act'] >= 0)
    assert both.sum() > n // 4
    if precision == 0:
        assert np.abs(V - ref['V']).max() / np.abs(ref['V']).max() <= 1e-9
    else:
        assert np.abs(V - ref['V']).max() <= MIXED_MV
        assert np.abs(act[both] - ref['act'][both]).max() <= MIXED_ACT_MS


def test_parameters_by_the_reference_names():
    """SetParameter names of the reference's tests (TestIonicModels.hpp:
    FR2000 ScaleFactorGks/Ito/Gkr, Mahajan and Maleckar ScaleFactor*; LR91's
    oxmeta conductance) find the same parameters as the library's names."""
    from chaste_amd import cmx
    assert cmx.find_param(cmx.FR2000, 'ScaleFactorGks') == 0
    assert cmx.find_param(cmx.FR2000, 'ScaleFactorIto') == 1
    assert cmx.find_param(cmx.FR2000, 'ScaleFactorGkr') == 2
    assert cmx.find_param(cmx.MAHAJAN2008, 'ScaleFactorGkr') == 2
    assert cmx.find_param(cmx.MALECKAR2008, 'ScaleFactorGna') == 0
    assert cmx.find_param(cmx.LR91, 'membrane_fast_sodium_current_conductance') == 0
    for model in (cmx.LR91, cmx.TT06_EPI, cmx.HH52, cmx.FOX2002, cmx.MAHAJAN2008, cmx.DN85, cmx.FR2000,
                  cmx.MALECKAR2008, cmx.N98, cmx.ORD_ENDO):
        names = [n for n, _ in cmx.param_info(model)]
        aliases = cmx.param_aliases(model)
        assert len(aliases) == len(names)
        for q, (n, a) in enumerate(zip(names, aliases)):
            assert cmx.find_param(model, n) == q
            if a:
                assert cmx.find_param(model, a) == q
    with pytest.raises(cmx.CmxError, match='No parameter named'):
        cmx.find_param(cmx.LR91, 'ScaleFactorGks')


# Maleckar 2008 has no adjoint: its forward-mode derivatives with respect to
# the dyadic calcium Ca_d grow past 1e60 through a strongly driven upstroke on
# the CPU as well, so its gradient cannot match a central difference.
ADJOINT_CASES = sorted(k for k in CASES if k != 'maleckar2008')


@pytest.mark.gpu
def test_adjoint_rejects_model_without_one():
    from chaste_amd import cmx
    print('case adjoint rejects maleckar2008', flush=True)
    coords, elems = cmx.slab_mesh(2, 4, 4, 0, 0.01)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.MALECKAR2008, cmx.RUSH_LARSEN,
                       dt_pde=0.01, dt_ode=0.01)
    obs = np.zeros(1, np.int32)
    with pytest.raises(cmx.CmxError, match='has no adjoint'):
        m.adjoint(0.0, 2, obs, 1, np.zeros((2, 1)))
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize('name', ADJOINT_CASES)
def test_adjoint_gradient_of_every_generated_model(name):
    """The discrete adjoint of the FP64 ICI step (cmx_adjoint) for each
    generated model with one, Rush-Larsen cells on a 2-D slab through a
    stimulated upstroke: the parameter gradient along a seeded direction
    against a central difference of the device's own forward runs (1e-4
    relative)."""
    from chaste_amd import cmx
    cid, _, _, _ = CASES[name]
    mag, dur, dt_ode = TISSUE[name]
    model = getattr(cmx, cid)
    print('case adjoint', name, flush=True)
    coords, elems = cmx.slab_mesh(2, 20, 6, 0, 0.01)
    n = coords.shape[0]
    stim = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), model, cmx.RUSH_LARSEN, stimuli=[(mag, 0.0, dur)],
                       stim_index=stim, dt_pde=0.01, dt_ode=dt_ode, ksp_rtol=1e-12)
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(model)])
    p0 = defaults[:, None] * np.ones((1, n))
    obs = np.arange(0, n, 7).astype(np.int32)
    steps, every = 200, 20
    data = np.full((steps // every, len(obs)), -20.0)

    def loss_at(p):
        m.set_state(y0)
        m.set_params(p)
        return m.adjoint(0.0, steps, obs, every, data, checkpoint_every=50)

    loss, g, _ = loss_at(p0)
    assert np.isfinite(loss) and np.isfinite(g).all()
    rng = np.random.default_rng(3)
    d = rng.standard_normal(p0.shape) * np.abs(p0) * 1e-5
    fd = (loss_at(p0 + d)[0] - loss_at(p0 - d)[0]) / 2
    ad = float((g * d).sum())
    m.close()
    print('%s adjoint: loss %.6e, directional %.6e, central difference %.6e, rel %.2e' % (
        name, loss, ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(fd) > 0
    assert abs(ad - fd) <= 1e-4 * abs(fd)
