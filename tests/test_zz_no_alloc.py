"""The allocation-free batched-cell entry point (cmx_cells_run_noalloc)
through the C ABI against cmx_cells_run on the same inputs: two calls in a
row on the static scratch array, the second with fewer cells, in FP64 and in
the mixed precision.  The caller's device buffers and stream come from the
HIP runtime libcmx.so itself is linked to, through ctypes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree / streams of libcmx.so's HIP runtime."""

    def __init__(self):
        from chaste_amd import cmx
        cmx.lib()                                   # loads /opt/rocm's libamdhip64 first
        self.L = ctypes.CDLL('libamdhip64.so.7')
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def check(self, rc):
        assert rc == 0, 'HIP error %d' % rc

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)))
        self.bufs.append(p)
        self.check(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))
        return p

    def download(self, p, shape):
        out = np.zeros(shape)
        self.check(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))
        return out

    def close(self):
        for p in self.bufs:
            self.check(self.L.hipFree(p))
        self.bufs = []


def _inputs(cmx, n, seed):
    rng = np.random.default_rng(seed)
    defaults = np.array([d for _, d in cmx.param_info(cmx.TT06_EPI)])
    params = defaults[:, None] * np.exp(rng.uniform(np.log(0.5), np.log(2.0), (len(defaults), n)))
    period = np.round(rng.uniform(300.0, 1000.0, n))
    t_end = np.round(rng.uniform(20.0, 400.0, n) / 0.005) * 0.005       # uneven run lengths
    y0 = np.repeat(np.asarray(cmx.initial_state(cmx.TT06_EPI))[:, None], n, axis=1)
    y0[0] += rng.uniform(-2.0, 2.0, n)
    stim = np.zeros((n, 5))
    stim[:, 0], stim[:, 1], stim[:, 2], stim[:, 3] = -80.0, 0.0, 2.0, period
    return t_end, stim, params, y0


def _noalloc(cmx, hip, model, precision, t_end, stim, params, y0, dt, thr):
    S = y0.shape[0]
    n = t_end.shape[0]
    d_y = hip.upload(np.full((S, n), -7.0))
    d_props = hip.upload(np.full((4, n), -7.0))
    c = cmx.CmxCells()
    c.model, c.integrator, c.precision, c.n_cells, c.dt = model, cmx.RUSH_LARSEN, precision, n, dt
    c.t_end = ctypes.cast(hip.upload(t_end), cmx._dp)
    c.stimuli = ctypes.cast(hip.upload(stim), ctypes.POINTER(cmx.CmxStimulus))
    c.n_params = params.shape[0]
    c.params = ctypes.cast(hip.upload(params), cmx._dp)
    c.threshold = thr
    stream = ctypes.c_void_p()
    hip.check(hip.L.hipStreamCreate(ctypes.byref(stream)))
    rc = cmx.lib().cmx_cells_run_noalloc(ctypes.byref(c), hip.upload(y0), d_y, d_props, stream)
    assert rc == 0, cmx.lib().cmx_cells_last_error()
    hip.check(hip.L.hipStreamSynchronize(stream))
    hip.check(hip.L.hipStreamDestroy(stream))
    y, props = hip.download(d_y, (S, n)), hip.download(d_props, (4, n))
    hip.close()
    return y, props


@pytest.mark.parametrize('precision', [0, 1], ids=['fp64', 'mixed'])
def test_noalloc_matches_cells_run(precision):
    from chaste_amd import cmx
    print('case no-alloc cells', precision, flush=True)
    hip = Hip()
    for n, seed in ((3000, 11), (1100, 12)):        # the second call is smaller
        t_end, stim, params, y0 = _inputs(cmx, n, seed)
        y_ref, p_ref, _ = cmx.cells_run(cmx.TT06_EPI, cmx.RUSH_LARSEN, t_end, 0.005,
                                        stimuli=[tuple(r[:4]) for r in stim], params=params, y0=y0,
                                        precision=precision, threshold=-30.0, lanes_per_cell=1)
        y, props = _noalloc(cmx, hip, cmx.TT06_EPI, precision, t_end, stim, params, y0, 0.005, -30.0)
        print('n %d: max |dy| %.3e, props equal %s, APs %d' % (n, np.abs(y - y_ref).max(),
                                                               np.array_equal(props, p_ref), int(p_ref[0].sum())))
        assert np.array_equal(y, y_ref)
        assert np.array_equal(props, p_ref)
        assert p_ref[0].sum() > 0


def test_noalloc_rejects_beyond_scratch_and_marks_bad_cells():
    from chaste_amd import cmx
    print('case no-alloc limits', flush=True)
    hip = Hip()
    c = cmx.CmxCells()
    c.model, c.integrator, c.precision, c.n_cells, c.dt = cmx.TT06_EPI, cmx.RUSH_LARSEN, 0, 118145, 0.005
    c.t_end = ctypes.cast(hip.upload(np.zeros(1)), cmx._dp)
    rc = cmx.lib().cmx_cells_run_noalloc(ctypes.byref(c), None, None, None, None)
    hip.close()
    assert rc == -1 and b'scratch' in cmx.lib().cmx_cells_last_error()
    # a t_end that is not a whole number of dt comes back as NaN; its
    # neighbours are unaffected
    t_end, stim, params, y0 = _inputs(cmx, 70, 3)
    t_end[5] += 0.0021
    y, props = _noalloc(cmx, hip, cmx.TT06_EPI, 0, t_end, stim, params, y0, 0.005, -30.0)
    assert np.isnan(y[:, 5]).all() and np.isnan(props[:, 5]).all()
    ok = np.arange(70) != 5
    assert np.isfinite(y[:, ok]).all()
