"""The O'Hara-Rudy 2011 endocardial cell (ORd), written from the published
equations (O'Hara et al., PLoS Comput Biol 7:e1002061, 2011).  The
reference tree holds no ORd model, code or data, so its arithmetic is
parity unpinned; what these tests pin:

* CPU (oracle): the published action potential -- paced at 1 Hz with the
  paper's -80 uA/uF, 0.5 ms stimulus from the published initial values,
  after 20 beats: resting potential near -88 mV, peak near +40 mV, APD90
  near 270 ms, maximum upstroke velocity near 260 V/s -- and rate
  dependence (APD90 shorter at a 500 ms cycle); the Rush-Larsen form against
  forward Euler (the clamped step converges to forward Euler's at first
  order; every state's error against forward Euler at dt 1e-6 halves with
  dt, ratio 2 +- 0.1, the check TestRushLarsen.hpp:179-196 makes for LR91).
* GPU: batched cells in FP64 against the oracle's single cells, FP32
  against FP64 (0.5 mV; onsets 0.1 ms, repolarisation 1 ms), and an ORd tissue against the oracle's run
  (1e-9 relative).
"""
import numpy as np
import pytest


def _features(v, dt, cl):
    """Resting potential, peak, APD90 (ms) and max dV/dt (V/s) of the last
    beat of a trace sampled every dt that starts with a stimulus every cl."""
    n = int(round(cl / dt))
    last = v[-(n + 1):]
    rest, peak = last[0], last.max()
    ipk = int(np.argmax(last))
    v90 = peak - 0.9 * (peak - rest)
    apd90 = (ipk + int(np.argmax(last[ipk:] < v90))) * dt
    return rest, peak, apd90, float((np.diff(last) / dt).max())


def test_ord_published_action_potential(oracle):
    M = oracle.ORD_ENDO
    y = oracle.model_initial_state(M)
    assert y.shape == (41,)
    v = oracle.cell_trace_v(M, y, 20000.0, 0.005, stim=(-80.0, 0.5, 0.0, 1000.0), integrator=oracle.RL)
    rest, peak, apd90, dvdt = _features(v, 0.005, 1000.0)
    print('ORd 1 Hz: rest %.2f peak %.2f APD90 %.1f dV/dt %.1f' % (rest, peak, apd90, dvdt))
    assert -89.0 <= rest <= -87.0
    assert 35.0 <= peak <= 45.0
    assert 255.0 <= apd90 <= 285.0
    assert 220.0 <= dvdt <= 300.0
    y = oracle.model_initial_state(M)
    v = oracle.cell_trace_v(M, y, 10000.0, 0.005, stim=(-80.0, 0.5, 0.0, 500.0), integrator=oracle.RL)
    apd500 = _features(v, 0.005, 500.0)[2]
    assert apd90 - 40.0 <= apd500 <= apd90 - 10.0


def test_ord_rush_larsen_against_forward_euler(oracle):
    M = oracle.ORD_ENDO
    y0 = oracle.model_initial_state(M)
    # clamped at the initial voltage for 1 ms: Rush-Larsen integrates the
    # gates exactly at fixed V, forward Euler to first order
    d = []
    for dt in (0.01, 0.005):
        a, b = y0.copy(), y0.copy()
        oracle.cell_compute_except_voltage(M, a, 0.0, 1.0, dt, integrator=oracle.RL)
        oracle.cell_compute_except_voltage(M, b, 0.0, 1.0, dt, integrator=oracle.FE)
        assert a[0] == y0[0]
        d.append(np.abs(a - b).max())
    assert 0 < d[0] <= 1e-4 and 1.9 <= d[0] / d[1] <= 2.1
    stim = (-80.0, 0.5, 0.0, -1.0)

    def run(integ, dt):
        y = y0.copy()
        oracle.cell_compute(M, y, 0.0, 1.0, dt, stim=stim, integrator=integ, steps_per_sample=int(round(1.0 / dt)))
        return y
    ref = run(oracle.FE, 1e-6)
    e1 = run(oracle.RL, 0.001) - ref
    e2 = run(oracle.RL, 0.0005) - ref
    assert ref[0] > -40.0                    # the upstroke is under way at 1 ms
    assert np.all(np.abs(e1 / e2 - 2.0) <= 0.1), e1 / e2


@pytest.mark.gpu
@pytest.mark.parametrize('integ', ['fe', 'rl'])
def test_ord_batched_cells_fp64_matches_oracle(oracle, integ):
    from chaste_amd import cmx
    print('case ord cells fp64', integ, flush=True)
    rng = np.random.default_rng(7)
    n = 96
    defaults = np.array([d for _, d in cmx.param_info(cmx.ORD_ENDO)])
    params = defaults[:, None] * np.exp(rng.uniform(np.log(0.5), np.log(2.0), (7, n)))
    period = np.round(rng.uniform(300.0, 1000.0, n))
    t_end = period * rng.integers(1, 3, n) + np.round(rng.uniform(0, 200, n))
    # forward Euler with the 2 ms stimulus is unstable: V reaches ~130 mV,
    # where the h_f gate's time constant is ~3e-4 ms; the paper's 0.5 ms
    # stimulus keeps V near +50 mV, where dt 0.002 is stable
    dt = 0.005 if integ == 'rl' else 0.002
    dur = 2.0 if integ == 'rl' else 0.5
    t_end = np.round(t_end / dt) * dt
    stimuli = [(-80.0, 0.0, dur, float(p)) for p in period]
    ig = cmx.RUSH_LARSEN if integ == 'rl' else cmx.FORWARD_EULER
    y, props, _ = cmx.cells_run(cmx.ORD_ENDO, ig, t_end, dt, stimuli=stimuli, params=params, precision=cmx.FP64,
                                threshold=-30.0)
    oi = oracle.RL if integ == 'rl' else oracle.FE
    scale = np.abs(np.asarray(cmx.initial_state(cmx.ORD_ENDO))) + 1e-3
    worst = 0.0
    for c in range(0, n, 7):
        yc = oracle.model_initial_state(oracle.ORD_ENDO)
        v = oracle.cell_trace_v(oracle.ORD_ENDO, yc, t_end[c], dt, stim=(-80.0, dur, 0.0, period[c]), integrator=oi,
                                params=params[:, c])
        worst = max(worst, float((np.abs(y[:, c] - yc) / scale).max()))
        pr = oracle.cell_properties(v, dt, -30.0)
        assert pr[0] == props[0][c] and pr[0] >= 1
        assert np.abs(pr[1:] - props[1:, c]).max() <= 1e-9
    print('ord cells fp64 worst scaled state error %.3e' % worst)
    assert worst <= 1e-9


@pytest.mark.gpu
def test_ord_batched_cells_mixed_against_fp64():
    from chaste_amd import cmx
    print('case ord cells mixed', flush=True)
    rng = np.random.default_rng(11)
    n = 512
    defaults = np.array([d for _, d in cmx.param_info(cmx.ORD_ENDO)])
    params = defaults[:, None] * np.exp(rng.uniform(np.log(0.5), np.log(2.0), (7, n)))
    period = np.round(rng.uniform(300.0, 1000.0, n))
    t_end = period * rng.integers(2, 5, n)
    stimuli = [(-80.0, 0.0, 2.0, float(p)) for p in period]
    out = {}
    for prec in (cmx.FP64, cmx.MIXED):
        out[prec] = cmx.cells_run(cmx.ORD_ENDO, cmx.RUSH_LARSEN, t_end, 0.005, stimuli=stimuli, params=params,
                                  precision=prec, threshold=-30.0)
    (y64, p64, _), (y32, p32, _) = out[cmx.FP64], out[cmx.MIXED]
    dv = np.abs(y32[0] - y64[0]).max()
    same = p32[0] == p64[0]
    print('ord mixed: max dV %.3e mV, AP counts equal %d / %d, onset err %.3e ms, end err %.3e ms' % (
        dv, same.sum(), n, np.abs(p32[1] - p64[1])[same].max(), np.abs(p32[3] - p64[3])[same].max()))
    assert dv <= 0.5
    assert same.all()
    assert np.abs(p32[1] - p64[1]).max() <= 0.1          # upstroke (onset) times
    assert np.abs(p32[3] - p64[3]).max() <= 1.0          # repolarisation times


@pytest.mark.gpu
@pytest.mark.parametrize('precision', [0, 1], ids=['fp64', 'mixed'])
def test_ord_tissue_matches_oracle(oracle, precision):
    from chaste_amd import cmx
    print('case ord tissue', precision, flush=True)
    dims = (2, 40, 10, 0)
    coords, elems = cmx.slab_mesh(*dims, 0.01)
    n = coords.shape[0]
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.ORD_ENDO, cmx.RUSH_LARSEN,
                       stimuli=[(-400000.0, 0.0, 0.5)], stim_index=stim_index, precision=precision,
                       dt_pde=0.01, dt_ode=0.005, ksp_rtol=1e-12 if precision == 0 else 1e-8, ksp_atol=0.0,
                       refine_sweeps=1 if precision else 0, use_graph=True)
    m.set_activation(0.0, 1, 0.0)
    m.step(0.0, 600)
    V, act = m.voltage(), m.activation()[4]
    ref = oracle.mono_run(*dims, model=oracle.ORD_ENDO, rush_larsen=True, sigma=(1.75, 1.75, 1.75),
                          stim_mag=-400000.0, stim_dur=0.5, stim_xlim=0.02, dt_ode=0.005, t_end=6.0, ksp_rtol=1e-12,
                          act_threshold=0.0)
    both = (act >= 0) & (ref['act'] >= 0)
    print('ord tissue: rel err %.3e, max dV %.3e mV, activated %d, act err %.3e' % (
        np.abs(V - ref['V']).max() / np.abs(ref['V']).max(), np.abs(V - ref['V']).max(), both.sum(),
        np.abs(act[both] - ref['act'][both]).max()))
    assert both.sum() > n // 2
    if precision == 0:
        assert np.abs(V - ref['V']).max() / np.abs(ref['V']).max() <= 1e-9
    else:
        assert np.abs(V - ref['V']).max() <= 0.5
        assert np.abs(act[both] - ref['act'][both]).max() <= 0.1
